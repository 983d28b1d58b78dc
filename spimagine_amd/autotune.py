"""Launch autotuner for the ray caster's frames and batches.

The kernels' launch knobs (spr_set_tuning) change speed, never results:
  slab_kb    per-wave LDS slab of the staged sampler (0 = global gathers only, 12, 16, 19)
  chunk      samples per staged box (0 = from the ray step: two 16-sample groups where the
             step allows 16; 16 = one group)
  staging    0 = each wave stages its own box, 1 = a workgroup's four waves stage one box
  tile_map   single frames: how 16x16 tiles are dealt to the XCDs (0 interleaved rows,
             3 = the same in descending estimated cost)
  wg_waves   single frames: waves per workgroup (0 = default: 1; 4)
  skip       empty-space skipping / early termination by brick bounds (0, 1)
The best values depend on voxel type, volume and frame size, step count, render mode and
the poses (how far the rays' footprints spread across a slab), so the tuner times every
candidate on a sample of the actual work and keeps the fastest per workload key, in memory
and optionally in a JSON file.  The context's tuning before the search (the library's
defaults unless the caller set others) stays unless a candidate beats it by more than
MARGIN: candidates within the timing noise of each other would otherwise trade places from
run to run, and a sample's near-tie can be a clear loss on one of its poses.  The reference has no counterpart: its OpenCL launch is a
fixed global size of one work item per pixel (spimagine/volumerender/volumerender.py:366-386).
"""
import itertools
import json
import os
import time

from . import _lib

# the batch kernel's persistent workgroups take tiles from per-XCD queues: tile_map and
# wg_waves do not apply
BATCH_CANDIDATES = tuple(
    [dict(slab_kb=0, chunk=0, staging=0), dict(slab_kb=12, chunk=0, staging=0)] +
    [dict(slab_kb=s, chunk=c, staging=st) for s, c, st in
     itertools.product((16, 19), (0, 16), (0, 1))])
# (skipping runs with the default 19 KiB slabs only)
FRAME_CANDIDATES = tuple(
    dict(tile_map=tm, wg_waves=ww, slab_kb=s, chunk=c, skip=sk)
    for tm, ww, (s, sk), c in itertools.product((0, 3), (0, 4), ((16, 0), (19, 0), (19, 1)),
                                                 (0, 16)))
CANDIDATES = BATCH_CANDIDATES  # the batch tuner's default
# a candidate replaces the incumbent tuning only when it is faster by more than this
# fraction (the spread of repeated timings of one candidate is ~0.5%)
MARGIN = 0.01

_cache = {}


def workload_key(ctx, jobs, w, h, params):
    """(voxel type, volume shape of the first job, frame, steps, filter, opacity mode)."""
    info = ctx.volume_info(jobs[0][0])
    return "%d:%s:%dx%d:%d:%d:%d" % (info["dtype"], "x".join(map(str, info["shape"])), w, h,
                                     params.max_steps, params.filter, params.alpha_pow != 0)


def _load_cache(cache_file):
    if cache_file and not _cache and os.path.exists(cache_file):
        try:
            _cache.update(json.load(open(cache_file)))
        except (OSError, ValueError):
            pass


def _store(key, best, table, cache_file):
    _cache[key] = {"best": best, "ms": table, "when": time.time()}
    if cache_file:
        d = os.path.dirname(cache_file)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1)


def _search(ctx, key, candidates, timer, cache_file, force, margin=MARGIN):
    """Times timer() under each candidate tuning; sets the fastest on ctx and returns
    (best, {candidate: ms}).  The candidate equal to ctx's tuning at the call (if one is)
    stays the best unless another is faster by more than `margin`."""
    _load_cache(cache_file)
    if key in _cache and not force:
        best = _cache[key]["best"]
        for k, v in best.items():
            ctx.set_tuning(k, v)
        return best, _cache[key]["ms"]
    before = {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS}
    table = {}
    try:
        for cand in candidates:
            for k, v in before.items():
                ctx.set_tuning(k, v)
            for k, v in cand.items():
                ctx.set_tuning(k, v)
            table[json.dumps(cand, sort_keys=True)] = timer()
    finally:
        for k, v in before.items():
            ctx.set_tuning(k, v)
    best = min(table, key=table.get)
    incumbent = next((c for c in table
                      if all(before.get(k) == v for k, v in json.loads(c).items())), None)
    if incumbent is not None and table[best] >= table[incumbent] * (1. - margin):
        best = incumbent
    best = json.loads(best)
    for k, v in best.items():
        ctx.set_tuning(k, v)
    _store(key, best, table, cache_file)
    return best, table


def tune_batch(ctx, jobs, w, h, params, candidates=BATCH_CANDIDATES, reps=2, cache_file=None,
               force=False):
    """Times ctx.render_batch(jobs, ...) under each candidate tuning, sets the fastest on
    ctx and returns (best, {candidate: ms}).  `jobs` should be a representative sample of
    the batch (a few volumes x poses); results are cached per workload key."""
    import torch
    stream = torch.cuda.Stream()

    def timer():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ctx.render_batch(jobs, w, h, stream=stream.cuda_stream, params=params)  # warm
        best = float("inf")
        for _ in range(reps):
            ev[0].record(stream)
            ctx.render_batch(jobs, w, h, stream=stream.cuda_stream, params=params)
            ev[1].record(stream)
            stream.synchronize()
            best = min(best, ev[0].elapsed_time(ev[1]))
        return best

    return _search(ctx, "batch:" + workload_key(ctx, jobs, w, h, params), candidates, timer,
                   cache_file, force)


def tune_frame(ctx, key, render, poses, stream, candidates=FRAME_CANDIDATES, reps=2,
               cache_file=None, force=False):
    """Single frames: render(k) enqueues one frame of sample pose k (0 <= k < poses) on
    `stream` (a torch.cuda.Stream) with the context's current tuning; each candidate is
    timed as the mean over the poses of the best of `reps` frames.  Sets the fastest on
    ctx; returns (best, {candidate: ms}).  key: the workload (see workload_key)."""
    import torch

    def timer():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        total = 0.
        for k in range(poses):
            render(k)  # warm
            best = float("inf")
            for _ in range(reps):
                ev[0].record(stream)
                render(k)
                ev[1].record(stream)
                stream.synchronize()
                best = min(best, ev[0].elapsed_time(ev[1]))
            total += best
        return total / poses

    return _search(ctx, "frame:" + key, candidates, timer, cache_file, force)
