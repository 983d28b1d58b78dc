"""Launch autotuner for batched max projections.

The kernels' launch knobs (spr_set_tuning) change speed, never results.  The batch
kernel's knob:
  slab_kb   per-wave LDS slab of the staged sampler (0 = global gathers only, 12, 16, 19)
(its persistent workgroups take tiles from per-XCD queues, so the single-frame knobs
tile_map and wg_waves do not apply).  The best value depends on voxel type, volume and
frame size, step count and the poses (how far the rays' footprints spread across a
slab), so the tuner times every candidate on a sample of the actual batch and keeps the
fastest for that workload key.  The
reference has no counterpart: its OpenCL launch is a fixed global size of one work item
per pixel (spimagine/volumerender/volumerender.py:366-386).
"""
import json
import os
import time

from . import _lib

CANDIDATES = tuple(dict(slab_kb=s) for s in (0, 12, 16, 19))

_cache = {}


def workload_key(ctx, jobs, w, h, params):
    """(voxel type, volume shape of the first job, frame, steps, filter, opacity mode)."""
    info = ctx.volume_info(jobs[0][0])
    return "%d:%s:%dx%d:%d:%d:%d" % (info["dtype"], "x".join(map(str, info["shape"])), w, h,
                                     params.max_steps, params.filter, params.alpha_pow != 0)


def _time_batch(ctx, jobs, w, h, params, stream, reps):
    import torch
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


def tune_batch(ctx, jobs, w, h, params, candidates=CANDIDATES, reps=2, cache_file=None,
               force=False):
    """Times ctx.render_batch(jobs, ...) under each candidate tuning, sets the fastest on
    ctx and returns (best, {candidate: ms}).  `jobs` should be a representative sample of
    the batch (a few volumes x poses); results are cached per workload key, in memory and,
    with cache_file, in a JSON file."""
    import torch
    key = workload_key(ctx, jobs, w, h, params)
    if cache_file and not _cache and os.path.exists(cache_file):
        try:
            _cache.update(json.load(open(cache_file)))
        except (OSError, ValueError):
            pass
    if key in _cache and not force:
        best = _cache[key]["best"]
        for k, v in best.items():
            ctx.set_tuning(k, v)
        return best, _cache[key]["ms"]
    before = {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS}
    stream = torch.cuda.Stream()
    table = {}
    try:
        for cand in candidates:
            for k, v in cand.items():
                ctx.set_tuning(k, v)
            table[json.dumps(cand, sort_keys=True)] = _time_batch(ctx, jobs, w, h, params,
                                                                  stream, reps)
    finally:
        for k, v in before.items():
            ctx.set_tuning(k, v)
    best = json.loads(min(table, key=table.get))
    for k, v in best.items():
        ctx.set_tuning(k, v)
    _cache[key] = {"best": best, "ms": table, "when": time.time()}
    if cache_file:
        d = os.path.dirname(cache_file)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1)
    return best, table
