"""Benchmarks of the MI355X volume ray caster.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]          # headline (default)
    python bench.py --workload batch [--timepoints 64] [--poses 16] [--steps 3]
    python bench.py --workload serve [--requests 400] [--qps 60]
    python bench.py --workload learned

headline  One step = one max-projection frame (one kernel launch) of the resident
          1024x1024x512 uint16 volume into 1024^2: 8000 seeded Gaussian blobs, background
          100, Poisson + read noise; perspective 45 deg, modelView
          translate(0,0,-4).rotY(30 deg), max_steps 1536 (each hitting ray takes
          (1536/16+1)*16 = 1552 samples, <= 1 voxel per step), min/max = data min/max,
          linear.  value = Gsamples/s over the timed steps, where the sample count is the
          reference's (rays hitting the box x 1552), independent of the implementation.
batch     64 timepoints of a 1024x1024x256 uint16 blob series (4000 blobs drifting by
          seeded random walks of 0.5 voxel per timepoint), all resident, x 16 poses = 1024
          frames of 1024^2 MIP + alpha.  One step = one replay of the HIP graph that holds
          the whole batch (one persistent launch); launch tuning from the autotuner.
serve     a Poisson request stream at fixed QPS against the render server: resident
          volumes from 256^3 float32 to 2048x2048x512 uint16, frames 512^2-2048^2, random
          poses, max projection / opacity / iso-surface; p50/p99 latency and frames/s.
learned   exact 1024^2 MIP of a held-out 512^3 uint16 volume vs the learned frame (two
          coarse 512^2 renders at max_steps/4 + the conv-net) over 8 seeded poses: frames/s
          of both, PSNR and max |error| per frame, and whether validation enables it.

Every line carries roofline.achieved = algorithmic bytes per launch (samples x 8 corners
x voxel bytes + the float32 output and alpha planes) / average kernel duration (HIP
events on the launch stream), over the 8 TB/s HBM peak, and cpu_baseline = the C oracle
(oracle/spimref.c, OpenMP) timed on a bounded block of rows of one frame of the workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gsamples/s + frames/s, MIP 1024x1024x512 u16 -> 1024^2; p99 frame latency"
HBM_PEAK_GBS = 8000.0
PMC_FILES = {"headline": os.path.join(ROOT, "profiles", "mip_headline_traffic.json"),
             "batch": os.path.join(ROOT, "profiles", "mip_batch_traffic.json")}
BOX = (-1, 1, -1, 1, -1, 1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def traffic_from_profile(name):
    path = PMC_FILES.get(name)
    if path and os.path.exists(path):
        try:
            return json.load(open(path)).get("traffic_bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def roofline(b_alg, kernel_ms, traffic):
    """achieved = algorithmic bytes / kernel time (the gather bandwidth the reference's
    sample count implies); with a measured traffic figure (bytes beyond L2 per launch,
    rocprofv3 counters) also the HBM bandwidth the launch actually drew."""
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "kernel_ms": round(kernel_ms, 4), "algorithmic_bytes": int(b_alg),
           "achieved_is": "effective gather bandwidth (reference samples x 8 corners)"}
    if traffic:
        measured = traffic / (kernel_ms * 1e-3) / 1e9
        out["measured_hbm_GBps"] = round(measured, 1)
        out["measured_hbm_frac"] = round(measured / HBM_PEAK_GBS, 4)
    return out


def cpu_threads():
    """Host threads of the CPU baselines: the CPUs this process may run on, at most 16 (one
    GPU's share of the host on an 8-GPU node; the sibling GPUs' host work uses the rest)."""
    return min(len(os.sched_getaffinity(0)), 16)


def cpu_baseline(vol, invP, invM, lo, hi, S, W, H, seconds, gpu_out, gpu_alpha, what):
    """The oracle on a bounded block of rows around the middle of the frame; returns the
    cpu_baseline record and the max |GPU - oracle| over those rows (inf if alpha differs)."""
    import oracle
    cores = cpu_threads()
    mid = H // 2
    per_ray = (S // 16 + 1) * 16

    def run_rows(r0, r1):
        t = time.perf_counter()
        o, a = oracle.max_project(vol, invP, invM, min_val=lo, max_val=hi, max_steps=S,
                                  width=W, height=H, rows=(r0, r1), nthreads=cores)
        return time.perf_counter() - t, o, a

    log("bench: cpu baseline (oracle, %d threads)" % cores)
    tcal, o, a = run_rows(mid, mid + cores)
    per_row = tcal / cores
    nrows = int(max(cores, min(H, (seconds / max(per_row, 1e-6)) // cores * cores)))
    r0 = max(0, mid - nrows // 2)
    r1 = min(H, r0 + nrows)
    tcpu, o, a = run_rows(r0, r1)
    reps = 1
    # a whole frame shorter than the budget: render it again until the budget is used
    while r1 - r0 == H and tcpu * (reps + 1) / reps <= seconds and reps < 64:
        tcpu += run_rows(r0, r1)[0]
        reps += 1
    hits = int((a[r0:r1] > 0).sum())
    cpu_samples = hits * per_ray * reps
    cpu = {"value": round(cpu_samples / tcpu / 1e9, 4), "unit": "Gsamples/s", "cores": cores,
           "kind": "port", "cpu": cpu_model(), "host_threads_visible": os.cpu_count(),
           "sample": "rows %d-%d of %s, %d time(s) (%d hit rays, %d samples), %.1f s"
                     % (r0, r1 - 1, what, reps, hits, cpu_samples, tcpu),
           "rows": [r0, r1 - 1], "reps": reps, "seconds": round(tcpu, 3)}
    parity = float(np.abs(gpu_out[r0:r1] - o[r0:r1]).max())
    if not np.array_equal(gpu_alpha[r0:r1], a[r0:r1]):
        parity = float("inf")
    return cpu, parity


def result_line(args, value, ms_per_step, dtype, config, **extra):
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gsamples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------------------ headline
def bench_headline(args):
    import torch
    from spimagine_amd import _lib, synth
    from spimagine_amd.volumerender.volumerender import render_matrices

    cfg = synth.CONFIGS["headline"]
    nz, ny, nx = cfg["shape"]
    W, H = cfg["out"]
    S = cfg["max_steps"]
    log("bench: generating %dx%dx%d uint16 blob volume on cuda:0" % (nx, ny, nz))
    t = time.time()
    vt = synth.config_volume("headline", device="cuda")
    torch.cuda.synchronize()
    log("bench: volume ready in %.1f s" % (time.time() - t))

    ctx = _lib.Context(0)
    tuning = {}
    for kv in args.tune:
        k, v = kv.split("=")
        ctx.set_tuning(k, int(v))
        tuning[k] = int(v)
    ctx.set_volume_device(vt.data_ptr(), np.uint16, cfg["shape"])
    lo, hi = ctx.volume_minmax()
    M, P = synth.headline_camera()
    invP, invM = render_matrices(M, P, (nx, ny, nz))

    def set_pose(iP, iM, filt=_lib.FILTER_LINEAR):
        ctx.set_params(iP, iM, BOX, lo, hi, 1., 0., .1, 21, 30, S, filt)

    set_pose(invP, invM)
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    alpha = torch.empty_like(out)
    torch.cuda.synchronize()
    # a dedicated (default-priority) stream: the frames and the events that time them are
    # enqueued on it (torch's default stream has handle 0, which the C ABI reads as
    # "the context's own stream")
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream
    assert sptr
    n_samp = ctx.count_samples(1, W, H)
    per_ray = (S // 16 + 1) * 16

    def frame():
        ctx.render_device(_lib.METHOD_MAX_PROJECT, 1, 0, W, H, out.data_ptr(), alpha.data_ptr(),
                          stream=sptr)

    # launch tuning: the autotuner on the headline pose and three of the seeded poses
    # (the mean frame time), unless --tune fixes it
    poses = synth.random_poses(8, seed=0)
    mats = [render_matrices(Mi, P, (nx, ny, nz)) for Mi in poses]
    if args.tune:
        tuned = {"source": "--tune"}
    else:
        from spimagine_amd import autotune
        sample = [(invP, invM)] + mats[:3]

        def render_k(k):
            set_pose(*sample[k])
            frame()

        log("bench: autotuning (%d candidates)" % len(autotune.FRAME_CANDIDATES))
        best, table = autotune.tune_frame(ctx, "u16:%dx%dx%d:%dx%d:%d" % (nx, ny, nz, W, H, S),
                                          render_k, len(sample), stream)
        tuned = {"source": "autotune", "best": best,
                 "candidates_ms": {k: round(v, 4) for k, v in table.items()}}
        set_pose(invP, invM)

    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        frame()
    torch.cuda.synchronize()

    log("bench: timing %d frames" % args.steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        frame()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = 1e3 * wall / args.steps
    gsps = n_samp * args.steps / wall / 1e9
    gpu_out = out.cpu().numpy()
    gpu_alpha = alpha.cpu().numpy()

    # per-frame latency over 8 seeded random poses
    lat_ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.latency_frames)]
    samples_pose = []
    for iP, iM in mats:
        set_pose(iP, iM)
        samples_pose.append(ctx.count_samples(1, W, H))
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(lat_ev):
        set_pose(*mats[i % len(mats)])
        a.record(stream)
        frame()
        b.record(stream)
    torch.cuda.synchronize()
    lat = [a.elapsed_time(b) for a, b in lat_ev] or [float("nan")]
    pose_gsps = float(np.mean([samples_pose[i % 8] / (lat[i] * 1e6) for i in range(len(lat_ev))]
                              or [float("nan")]))
    set_pose(invP, invM)
    # empty-space skipping off and on (SPR_TUNE_SKIP), the headline pose and the 8 poses
    skipping = skip_table(ctx, stream, frame, lambda k: set_pose(*([(invP, invM)] + mats)[k]),
                          [n_samp] + samples_pose, args.skip_frames)
    # the nearest filter (interpolation="nearest"): staged as the linear filter, or every
    # sample gathered from global memory (slab_kb 0)
    nearest = filter_table(ctx, stream, frame,
                           lambda k: set_pose(*([(invP, invM)] + mats)[k], _lib.FILTER_NEAREST),
                           [n_samp] + samples_pose, args.skip_frames)
    set_pose(invP, invM)

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib << 19  # int16 elements
        src = torch.ones(n, dtype=torch.int16, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    b_alg = n_samp * 8 * 2 + W * H * 4 * 2
    cpu = parity = None
    vol = synth.device_volume_to_numpy(vt, np.uint16)
    upload = measure_upload(ctx, vol)
    host_frame = measure_host_frames(ctx, W, H, 10)
    if not args.no_cpu:
        cpu, parity = cpu_baseline(vol, invP, invM, lo, hi, S, W, H, args.cpu_seconds,
                                   gpu_out, gpu_alpha, "the same 1024^2 frame")
    result_line(
        args, gsps, ms_per_step, "f32",
        {"workload": "MIP 1024x1024x512 uint16 volume (8000 seeded blobs, Poisson + read noise) "
                     "-> 1024x1024 float32, max_steps 1536, perspective 45 deg, "
                     "translate(0,0,-4).rotY(30 deg), linear sampler",
         "voxel": "u16", "frame": [W, H], "max_steps": S, "samples_per_frame": int(n_samp),
         "samples_per_hit_ray": per_ray,
         "tuning": {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS},
         "tuning_source": tuned},
        frames_per_s=round(1e3 / ms_per_step, 2),
        latency_ms={"p50": round(pct(lat, 50), 4), "p99": round(pct(lat, 99), 4),
                    "frames": len(lat_ev), "poses": 8, "gsamples_per_s": round(pose_gsps, 3)},
        skipping=skipping, nearest_filter=nearest,
        roofline=roofline(b_alg, kernel_ms, traffic_from_profile("headline")),
        cpu_baseline=cpu, parity_max_abs_diff_vs_oracle=parity, upload=upload,
        host_frame_ms=host_frame)


def skip_table(ctx, stream, frame, set_pose, samples, reps):
    """Per-pose frame times (HIP events on the frame stream) with SPR_TUNE_SKIP 0 and 1:
    pose 0 is the headline pose, the rest the seeded random poses.  Restores the context's
    skip setting.  Samples are the reference count either way: skipping raises the
    effective rate, not the work.  reps 0: no table (None)."""
    import torch
    if reps <= 0:
        return None
    keep = ctx.get_tuning("skip")
    out = {}
    for skip in (0, 1):
        ctx.set_tuning("skip", skip)
        ms = []
        for k in range(len(samples)):
            set_pose(k)
            frame()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                frame()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1) / reps)
        out["on" if skip else "off"] = {
            "headline_ms": round(ms[0], 4),
            "headline_gsamples_per_s": round(samples[0] / (ms[0] * 1e6), 2),
            "poses_mean_ms": round(float(np.mean(ms[1:])), 4),
            "poses_gsamples_per_s": round(float(np.mean([n / (t * 1e6) for n, t in
                                                         zip(samples[1:], ms[1:])])), 2),
            "poses_ms": [round(t, 4) for t in ms[1:]]}
    ctx.set_tuning("skip", keep)
    return out


def filter_table(ctx, stream, frame, set_pose, samples, reps):
    """Per-pose frame times of the nearest filter (set_pose(k) selects it) staged in LDS
    slabs (the context's slab_kb) and with every sample gathered from global memory
    (slab_kb 0).  Restores slab_kb.  reps 0: no table (None)."""
    import torch
    if reps <= 0:
        return None
    keep = ctx.get_tuning("slab_kb")
    out = {}
    for name, slab in (("staged", keep), ("global", 0)):
        ctx.set_tuning("slab_kb", slab)
        ms = []
        for k in range(len(samples)):
            set_pose(k)
            frame()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                frame()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1) / reps)
        out[name] = {"headline_ms": round(ms[0], 4),
                     "headline_gsamples_per_s": round(samples[0] / (ms[0] * 1e6), 2),
                     "poses_mean_ms": round(float(np.mean(ms[1:])), 4)}
    ctx.set_tuning("slab_kb", keep)
    return out


def measure_host_frames(ctx, W, H, n):
    """The drop-in host path (spr_render, as VolumeRenderer.render calls it): one frame of
    the context's current parameters rendered and its three planes (output, alpha, depth)
    copied into fresh host arrays, wall time per call (median of n).  Pageable numpy arrays
    (what VolumeRenderer allocates) and, beside them, arrays in pinned host memory."""
    import torch
    from spimagine_amd import _lib
    res = {}
    for kind in ("pageable", "pinned"):
        ts = []
        for i in range(n + 2):
            if kind == "pageable":
                planes = [np.empty((H, W), np.float32) for _ in range(3)]
            else:
                planes = [torch.empty((H, W), pin_memory=True).numpy() for _ in range(3)]
            t0 = time.perf_counter()
            ctx.render(_lib.METHOD_MAX_PROJECT, 1, 0, W, H, *planes)
            ts.append(time.perf_counter() - t0)
        res[kind + "_ms"] = round(1e3 * float(np.median(ts[2:])), 3)
    return res


def measure_upload(ctx, vol):
    """The format path on the same volume, outside the timed frames: host (pageable numpy)
    -> device upload + padded-layout build (spr_volume_add), and the GPU min/max."""
    import torch
    t0 = time.perf_counter()
    slot = ctx.volume_add(vol)
    t1 = time.perf_counter()
    ctx.volume_release(slot)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    lo, hi = ctx.volume_minmax()
    t3 = time.perf_counter()
    return {"bytes": int(vol.nbytes), "host_to_padded_ms": round(1e3 * (t1 - t0), 2),
            "host_to_padded_GBps": round(vol.nbytes / (t1 - t0) / 1e9, 2),
            "minmax_ms": round(1e3 * (t3 - t2), 3), "minmax": [lo, hi]}


# ------------------------------------------------------------------------------ iso
def iso_reference_samples(depth, tnear, invP, invM, W, H, S, rows, n_box_hits):
    """Samples the reference's iso march takes for the frame rows [r0, r1) (iso_kernel.cl:
    83-196): a ray that crosses the iso value at step i takes i + 1 march samples, j <= 10
    refinement samples and 24 gradient samples; a ray inside the box that never crosses
    takes S.  i and j come from the outputs, t_hit = tnear + i dt + j dt/10 with dt from
    the float64 box intersection (exact to within one sample per ray); n_box_hits: the rays
    of those rows that enter the box."""
    r0, r1 = rows
    iP = np.asarray(invP, np.float64).reshape(4, 4)
    iM = np.asarray(invM, np.float64).reshape(4, 4)
    ys, xs = np.mgrid[r0:r1, 0:W]
    hit = np.isfinite(depth[r0:r1])
    u, v = xs[hit] / W * 2 - 1, ys[hit] / H * 2 - 1
    one = np.ones_like(u)
    front = np.stack([u, v, -one, one], -1) @ iP.T
    front /= front[:, 3:]
    back = np.stack([u, v, one, one], -1) @ iP.T
    back /= back[:, 3:]
    orig = front @ iM.T
    orig /= orig[:, 3:]
    d = back - front
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    direc = (d @ iM.T)[:, :3]
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (-1 - orig[:, :3]) / direc
        t1 = (1 - orig[:, :3]) / direc
    tfar = np.nanmin(np.maximum(t0, t1), axis=1)
    tn = tnear[r0:r1][hit].astype(np.float64)
    dt = (tfar - tn) / (S - 1)
    x = (depth[r0:r1][hit].astype(np.float64) - tn) / dt
    i = np.clip(np.ceil(x - 1e-6) - 1, 1, S - 1)
    j = np.clip(np.rint((x - i) * 10), 1, 10)
    n_hit = int(hit.sum())
    return float(np.sum(i + 1 + j + 24) + (n_box_hits - n_hit) * S)


def bench_iso(args):
    """Iso surface (full pipeline) and opacity-mode frames of the 512^3 uint16 volume into
    1024^2 over 8 seeded poses, max_steps 896, skipping off and on."""
    import torch
    from spimagine_amd import _lib, synth
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices

    cfg = synth.CONFIGS["iso"]
    nz, ny, nx = cfg["shape"]
    W, H = cfg["out"]
    S = cfg["max_steps"]
    log("bench: generating 512^3 uint16 blob volume on cuda:0")
    vt = synth.config_volume("iso", device="cuda")
    ctx = _lib.Context(0)
    for kv in args.tune:  # launch settings fixed by hand (default: the library's)
        k, v = kv.split("=")
        ctx.set_tuning(k, int(v))
    ctx.set_volume_device(vt.data_ptr(), np.uint16, cfg["shape"])
    lo, hi = ctx.volume_minmax()
    iso_max = 2 * (lo + .3 * (hi - lo))
    P = mat4_perspective(45, 1., .1, 10)
    mats = [render_matrices(M, P, (nx, ny, nz)) for M in synth.random_poses(8, seed=0)]
    bufs = {k: torch.empty((H, W), device="cuda") for k in ("out", "alpha", "depth", "occ")}
    normals = torch.empty((H, W, 3), device="cuda")
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream

    def set_pose(k, mode, steps=S):
        iP, iM = mats[k]
        if mode == "iso":
            ctx.set_params(iP, iM, BOX, lo, iso_max, 1., 0., .1, 21, 30, steps, _lib.FILTER_LINEAR)
        else:
            ctx.set_params(iP, iM, BOX, lo, hi, 1., .6 if mode == "alpha" else 0., .1, 21, 30,
                           steps, _lib.FILTER_LINEAR)

    def frame(mode):
        if mode == "iso":
            ctx.render_device(_lib.METHOD_ISO_SURFACE, 1, 0, W, H, bufs["out"].data_ptr(),
                              bufs["alpha"].data_ptr(), bufs["depth"].data_ptr(),
                              normals.data_ptr(), bufs["occ"].data_ptr(), stream=sptr)
        else:
            ctx.render_device(_lib.METHOD_MAX_PROJECT, 1, 0, W, H, bufs["out"].data_ptr(),
                              bufs["alpha"].data_ptr(), stream=sptr)

    def pose_ms(mode, k, reps):
        set_pose(k, mode)
        frame(mode)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            frame(mode)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # reference iso samples per pose (from the outputs) and rays entering the box
    iso_samples, box_hits = [], []
    for k in range(8):
        set_pose(k, "iso")
        rays = ctx.count_samples(1, W, H) // ((S // 16 + 1) * 16)
        frame("iso")
        torch.cuda.synchronize()
        d, a = bufs["depth"].cpu().numpy(), bufs["alpha"].cpu().numpy()
        box_hits.append(rays)
        iso_samples.append(iso_reference_samples(d, a, mats[k][0], mats[k][1], W, H, S, (0, H),
                                                 rays))
    table = {}
    keep = ctx.get_tuning("skip")
    for skip in (0, 1):
        ctx.set_tuning("skip", skip)
        for mode in ("iso", "alpha"):
            ms = [pose_ms(mode, k, args.skip_frames) for k in range(8)]
            table.setdefault(mode, {})["on" if skip else "off"] = {
                "poses_ms": [round(t, 4) for t in ms], "mean_ms": round(float(np.mean(ms)), 4)}
    ctx.set_tuning("skip", keep)
    # the same poses at the drop-in renderer's default step count (max_steps 200, the
    # reference's config default): max projection, opacity mode and iso frame
    default_steps = {}
    for mode in ("mip", "alpha", "iso"):
        ms = []
        for k in range(8):
            set_pose(k, mode, 200)
            frame(mode)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(max(args.skip_frames, 1)):
                frame(mode)
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1) / max(args.skip_frames, 1))
        default_steps[mode] = {"mean_ms": round(float(np.mean(ms)), 4),
                               "poses_ms": [round(t, 4) for t in ms]}
    log("bench: timing %d iso frames (8 poses, skipping %s)" % (args.steps, "on" if keep else "off"))
    for k in range(8):
        set_pose(k, "iso")
        frame("iso")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        set_pose(i % 8, "iso")
        frame("iso")
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = 1e3 * wall / args.steps
    samp = float(np.mean([iso_samples[i % 8] for i in range(args.steps)]))
    b_alg = samp * 8 * 2 + W * H * 4 * (3 + 3)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_iso(synth.device_volume_to_numpy(vt, np.uint16), mats[0], iso_max / 2.,
                               S, W, H, args.cpu_seconds)
    print(json.dumps({
        "metric": "iso-surface + opacity frames/s, 512^3 u16 -> 1024^2, 8 poses",
        "value": round(1e3 / ms_per_step, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 8, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "iso-surface (Phong + normal blur + occlusion + shading) of a "
                               "512^3 uint16 volume (2000 seeded blobs, Poisson + read noise) "
                               "-> 1024x1024, iso at 0.3 of the data range, max_steps 896, "
                               "8 seeded poses; opacity mode at alpha_pow 0.6 timed beside it, "
                               "and all three modes at max_steps 200 (max_steps_200_ms)",
                   "voxel": "u16", "frame": [W, H], "max_steps": S,
                   "tuning": {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS}},
        "frames_per_s": round(1e3 / ms_per_step, 2),
        "iso_reference_samples_per_frame": [int(v) for v in iso_samples],
        "gsamples_per_s": round(samp / (kernel_ms * 1e6), 2),
        "skipping_ms": table,
        "max_steps_200_ms": default_steps,
        "roofline": dict(roofline(b_alg, kernel_ms, None),
                         kernel="iso pipeline per frame (march k_iso_staged + 6 post-pass "
                                "kernels; the march dominates, see profiles/)",
                         frac_skipping_off=round(
                             float(np.mean([b / (t * 1e-3) for b, t in zip(
                                 [v * 16 + W * H * 24 for v in iso_samples],
                                 table["iso"]["off"]["poses_ms"])])) / 1e9 / HBM_PEAK_GBS, 4)
                         if table else None),
        "cpu_baseline": cpu}), flush=True)


def cpu_baseline_iso(vol, mat, iso_val, S, W, H, seconds):
    """The oracle's iso march on a block of rows around the middle of the frame, in
    reference samples per second (counted as iso_reference_samples does)."""
    import oracle
    cores = cpu_threads()
    invP, invM = mat
    mid = H // 2
    t = time.perf_counter()
    oracle.iso_surface(vol, invP, invM, iso_val=np.float32(iso_val), max_steps=S, width=W,
                       height=H, rows=(mid, mid + cores), nthreads=cores)
    per_row = (time.perf_counter() - t) / cores
    n = int(max(cores, min(H, seconds / max(per_row, 1e-6) // cores * cores)))
    r0 = max(0, mid - n // 2)
    r1 = min(H, r0 + n)
    t = time.perf_counter()
    _, a, d, _ = oracle.iso_surface(vol, invP, invM, iso_val=np.float32(iso_val), max_steps=S,
                                    width=W, height=H, rows=(r0, r1), nthreads=cores)
    tcpu = time.perf_counter() - t
    reps = 1
    # a whole frame shorter than the budget: render it again until the budget is used
    while r1 - r0 == H and tcpu * (reps + 1) / reps <= seconds and reps < 64:
        t = time.perf_counter()
        oracle.iso_surface(vol, invP, invM, iso_val=np.float32(iso_val), max_steps=S, width=W,
                           height=H, rows=(r0, r1), nthreads=cores)
        tcpu += time.perf_counter() - t
        reps += 1
    # rays of those rows that enter the box: iso misses and box misses both read alpha 0, so
    # count the box hits with the max projection's sample counter geometry on the host
    iP = np.asarray(invP, np.float64).reshape(4, 4)
    iM = np.asarray(invM, np.float64).reshape(4, 4)
    ys, xs = np.mgrid[r0:r1, 0:W]
    u, v = xs / W * 2 - 1, ys / H * 2 - 1
    one = np.ones_like(u)
    front = np.stack([u, v, -one, one], -1) @ iP.T
    front /= front[..., 3:]
    back = np.stack([u, v, one, one], -1) @ iP.T
    back /= back[..., 3:]
    orig = (front @ iM.T)
    orig = orig[..., :3] / orig[..., 3:]
    dd = back - front
    dd /= np.linalg.norm(dd, axis=-1, keepdims=True)
    direc = (dd @ iM.T)[..., :3]
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (-1 - orig) / direc
        t1 = (1 - orig) / direc
    tmin = np.nanmax(np.minimum(t0, t1), axis=-1)
    tmax = np.nanmin(np.maximum(t0, t1), axis=-1)
    rays = int((tmax > tmin).sum())
    samples = iso_reference_samples(d, a, invP, invM, W, H, S, (r0, r1), rays)
    return {"value": round(samples * reps / tcpu / 1e9, 4), "unit": "Gsamples/s",
            "cores": cores, "kind": "port", "cpu": cpu_model(),
            "host_threads_visible": os.cpu_count(),
            "sample": "iso march (oracle) of rows %d-%d of pose 0, %d time(s) (%d rays in the "
                      "box, ~%d reference samples each), %.1f s"
                      % (r0, r1 - 1, reps, rays, samples, tcpu)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# ------------------------------------------------------------------------------ batch
def bench_batch(args):
    import torch
    from spimagine_amd import _lib, autotune, synth
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices

    shape = (256, 1024, 1024)
    nz, ny, nx = shape
    W = H = args.frame
    T, NP = args.timepoints, args.poses
    S = synth.steps_for_shape((nx, ny, nz))
    ctx = _lib.Context(0)
    log("bench: generating %d timepoints of %dx%dx%d uint16 on cuda:0" % (T, nx, ny, nz))
    t = time.time()
    slots = []
    lo = hi = None
    host_t0 = None
    for k, vt in enumerate(synth.blob_series_device(shape, 4000, (3., 10.), (500., 3000.), 100.,
                                                    True, 5., T, drift=.5, seed=0)):
        if k == 0:
            u = vt.to(torch.int32) & 0xFFFF
            lo, hi = float(u.min()), float(u.max())
            del u
            ctx.set_volume_device(vt.data_ptr(), np.uint16, shape)  # slot 0: sample counts
            if not args.no_cpu:
                host_t0 = synth.device_volume_to_numpy(vt, np.uint16)
        slots.append(ctx.volume_add(dev_ptr=vt.data_ptr(), dtype=np.uint16, shape=shape))
        del vt
    torch.cuda.synchronize()
    log("bench: %d volumes resident in %.1f s" % (T, time.time() - t))

    P = mat4_perspective(45, W / float(H), .1, 10)
    poses = synth.random_poses(NP, seed=1)
    mats = [render_matrices(Mi, P, (nx, ny, nz)) for Mi in poses]
    planes = torch.empty((T * NP, 2, H, W), dtype=torch.float32, device="cuda")
    jobs = []
    for ti in range(T):
        for pi in range(NP):
            f = ti * NP + pi
            jobs.append((slots[ti], mats[pi][0], mats[pi][1], planes[f, 0].data_ptr(),
                         planes[f, 1].data_ptr()))
    samples_pose = []
    for iP, iM in mats:
        ctx.set_params(iP, iM, BOX, lo, hi, 1., 0., .1, 21, 30, S, _lib.FILTER_LINEAR)
        samples_pose.append(ctx.count_samples(1, W, H))
    n_samp = T * sum(samples_pose)
    params = _lib.make_params(mats[0][0], mats[0][1], BOX, lo, hi, 1., 0., .1, 21, 30, S,
                              _lib.FILTER_LINEAR)
    ctx.set_params(mats[0][0], mats[0][1], BOX, lo, hi, 1., 0., .1, 21, 30, S, _lib.FILTER_LINEAR)

    if args.tune:
        for kv in args.tune:
            k, v = kv.split("=")
            ctx.set_tuning(k, int(v))
        tuned = {"source": "--tune"}
    else:
        sample = [jobs[ti * NP + pi] for ti in (0, T // 2) for pi in range(0, NP, max(1, NP // 4))]
        log("bench: autotuning on %d frames" % len(sample))
        best, table = autotune.tune_batch(ctx, sample, W, H, params)
        tuned = {"source": "autotune", "candidates_ms": table}
    tuning = {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS}
    log("bench: tuning %s" % tuning)

    graph = ctx.batch_graph(jobs, W, H)
    stream = torch.cuda.Stream()
    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        graph.launch(stream.cuda_stream)
    stream.synchronize()
    log("bench: timing %d replays of %d frames" % (args.steps, len(jobs)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        graph.launch(stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = 1e3 * wall / args.steps

    # one frame of the batch against the same frame rendered alone
    f = (T // 3) * NP + NP // 2
    single = torch.empty((2, H, W), dtype=torch.float32, device="cuda")
    p1 = _lib.make_params(jobs[f][1], jobs[f][2], BOX, lo, hi, 1., 0., .1, 21, 30, S,
                          _lib.FILTER_LINEAR)
    ctx.render_slot(jobs[f][0], p1, _lib.METHOD_MAX_PROJECT, W, H, single[0].data_ptr(),
                    single[1].data_ptr(), stream=stream.cuda_stream)
    stream.synchronize()
    same_alone = bool(torch.equal(single, planes[f]))

    b_alg = n_samp * 8 * 2 + len(jobs) * W * H * 4 * 2
    cpu = parity = None
    if not args.no_cpu:
        g0 = planes[0].cpu().numpy()
        cpu, parity = cpu_baseline(host_t0, mats[0][0], mats[0][1], lo, hi, S, W, H,
                                   args.cpu_seconds, g0[0], g0[1], "frame 0 (timepoint 0, pose 0)")
    graph.close()
    result_line(
        args, n_samp * args.steps / wall / 1e9, ms_per_step, "f32",
        {"workload": "batched MIP + alpha: %d timepoints of a 1024x1024x256 uint16 blob series "
                     "(4000 blobs, seeded random-walk drift 0.5 voxel/timepoint, Poisson + read "
                     "noise) x %d poses -> %d frames of %dx%d float32, max_steps %d, one HIP "
                     "graph replay per step" % (T, NP, len(jobs), W, H, S),
         "voxel": "u16", "frames_per_step": len(jobs), "frame": [W, H], "max_steps": S,
         "samples_per_step": int(n_samp), "resident_volume_bytes":
             int(sum(ctx.volume_info(s)["bytes"] for s in slots)),
         "tuning": tuning, "tuning_source": tuned},
        frames_per_s=round(len(jobs) * 1e3 / ms_per_step, 2),
        roofline=roofline(b_alg, kernel_ms, traffic_from_profile("batch")),
        cpu_baseline=cpu, parity_max_abs_diff_vs_oracle=parity,
        batch_frame_equals_single_render=same_alone)


# ------------------------------------------------------------------------------ serve
def bench_serve(args):
    import torch
    from spimagine_amd import _lib, synth
    from spimagine_amd.serving import RenderServer, Request, poisson_load
    from spimagine_amd.utils import mat4_perspective

    budget = int(args.hbm_budget_gib * (1 << 30)) if args.hbm_budget_gib else None
    server = RenderServer(0, hbm_budget=budget, max_batch=args.max_batch, workers=args.workers)
    info = {}
    t = time.time()
    for k, (name, shape, dt, nb) in enumerate(synth.SERVING_VOLUMES):
        vol = synth.serving_volume(shape, dt, nb, k)
        if dt == np.float32:
            rng = (0., float(vol.max()))
        else:
            rng = (float(vol.min()), float(vol.max()))
        server.register(name, vol)
        info[name] = dict(shape=shape, dtype=np.dtype(dt).name, range=rng,
                          steps=synth.steps_for_shape(shape[::-1]))
    log("bench: %d volumes registered in %.1f s" % (len(info), time.time() - t))
    rng = np.random.default_rng(5)
    sizes = [512, 1024, 2048]
    names = sorted(info)
    reqs = []
    for i in range(args.requests):
        name = names[rng.integers(len(names))]
        d = info[name]
        side = sizes[rng.integers(len(sizes))]
        M = synth.random_poses(1, seed=int(rng.integers(1 << 30)))[0]
        P = mat4_perspective(45, 1., .1, 10)
        mode = rng.choice(["max_project"] * 8 + ["alpha", "iso_surface"])
        lo, hi = d["range"]
        if mode == "iso_surface":
            r = Request(name, M, P, size=(side, side), method="iso_surface", min_val=lo,
                        max_val=2 * (lo + .3 * (hi - lo)), max_steps=d["steps"])
        else:
            r = Request(name, M, P, size=(side, side), min_val=lo, max_val=hi,
                        alpha_pow=.6 if mode == "alpha" else 0., max_steps=d["steps"])
        reqs.append(r)
    # warm: every volume admitted and every kernel variant loaded once
    for r in reqs[:args.warmup]:
        server.render_alone(r)
    torch.cuda.synchronize()
    log("bench: %d requests at %.0f QPS" % (len(reqs), args.qps))
    t0 = time.perf_counter()
    futs = poisson_load(server, reqs, args.qps, seed=6)
    res = [f.result() for f in futs]
    wall = time.perf_counter() - t0
    lat = [r.latency_s * 1e3 for r in res]
    batch = [r.batch_size for r in res]
    # the tail by request kind: (volume, frame side, method) -> latency p50 / p99 / max and
    # the mean time from dispatch to result (latency minus queueing)
    kinds = {}
    for q, r in zip(reqs, res):
        m = "alpha" if q.settings["alpha_pow"] else q.method
        kinds.setdefault("%s %d %s" % (q.volume, q.size[0], m), []).append(
            (r.latency_s * 1e3, (r.latency_s - r.queued_s) * 1e3))
    by_kind = {k: {"n": len(v), "p50": round(pct([a for a, _ in v], 50), 2),
                   "p99": round(pct([a for a, _ in v], 99), 2),
                   "max": round(max(a for a, _ in v), 2),
                   "service_mean": round(float(np.mean([b for _, b in v])), 2)}
               for k, v in sorted(kinds.items())}
    px = sum(r.size[0] * r.size[1] for r in reqs)
    server.close()
    # algorithmic bytes of the max-projection requests (reference samples of the rays that
    # enter the box, from a float64 box test, x 8 corners x voxel bytes, plus the two output
    # planes), over the whole run's wall time: idle gaps between arrivals included
    from spimagine_amd.volumerender.volumerender import render_matrices
    b_alg, mip = 0., []
    for r in reqs:
        if r.method != "max_project":
            continue
        d = info[r.volume]
        nz, ny, nx = d["shape"]
        iP, iM = render_matrices(r.model_view, r.projection, (nx, ny, nz))
        rays = box_hit_rays(iP, iM, r.size[0], r.size[1])
        per_ray = (r.settings["max_steps"] // 16 + 1) * 16
        b_alg += rays * per_ray * 8 * np.dtype(d["dtype"]).itemsize + r.size[0] * r.size[1] * 8
        mip.append((r, iP, iM))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_requests(server.volumes, mip[:4], args.cpu_seconds)
    print(json.dumps({
        "metric": "serving: p50/p99 frame latency and frames/s, Poisson arrivals over mixed "
                  "resident volumes", "value": round(len(res) / wall, 2), "unit": "frames/s",
        "n_gpus": 1, "steps": len(reqs), "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / len(reqs), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Poisson arrivals at %.0f QPS, %d requests over %d resident "
                               "volumes (256^3 f32 .. 2048x2048x512 u16), frames 512^2-2048^2, "
                               "80%% MIP / 10%% opacity / 10%% iso" % (args.qps, len(reqs), len(info)),
                   "volumes": {k: {"shape": list(v["shape"]), "dtype": v["dtype"]}
                               for k, v in info.items()},
                   "max_batch": args.max_batch, "workers": args.workers,
                   "hbm_budget_bytes": server.hbm_budget},
        "latency_ms": {"p50": round(pct(lat, 50), 3), "p99": round(pct(lat, 99), 3),
                       "mean": round(float(np.mean(lat)), 3)},
        "latency_by_kind_ms": by_kind,
        "mean_batch": round(float(np.mean(batch)), 2), "gpixels_per_s": round(px / wall / 1e9, 4),
        "batches": server.batches,
        "roofline": dict(roofline(b_alg, 1e3 * wall, None),
                         kernel="all launches of the run: max-projection requests' algorithmic "
                                "bytes over the run's wall time (arrival gaps included)"),
        "cpu_baseline": cpu}), flush=True)


def box_hit_rays(invP, invM, W, H, box=BOX):
    """Rays of a W x H frame that enter the box (ray set-up of volume_kernel.cl:48-93 in
    float64)."""
    iP = np.asarray(invP, np.float64).reshape(4, 4)
    iM = np.asarray(invM, np.float64).reshape(4, 4)
    ys, xs = np.mgrid[0:H, 0:W]
    u, v = xs / W * 2 - 1, ys / H * 2 - 1
    one = np.ones_like(u)
    front = np.stack([u, v, -one, one], -1) @ iP.T
    front /= front[..., 3:]
    back = np.stack([u, v, one, one], -1) @ iP.T
    back /= back[..., 3:]
    orig = front @ iM.T
    orig = orig[..., :3] / orig[..., 3:]
    d = back - front
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    direc = (d @ iM.T)[..., :3]
    lo = np.array(box[0::2], np.float64)
    hi = np.array(box[1::2], np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - orig) / direc
        t1 = (hi - orig) / direc
    tmin = np.nanmax(np.minimum(t0, t1), axis=-1)
    tmax = np.nanmin(np.maximum(t0, t1), axis=-1)
    return int((tmax > tmin).sum())


def cpu_baseline_requests(volumes, mip, seconds):
    """The oracle on a bounded block of rows of each of a few max-projection requests of the
    run; Gsamples/s over all of them."""
    import oracle
    cores = cpu_threads()
    per = seconds / max(1, len(mip))
    total_s = total_t = 0.
    parts = []
    for r, iP, iM in mip:
        vol = volumes[r.volume].data
        W, H = r.size
        st = r.settings
        per_ray = (st["max_steps"] // 16 + 1) * 16
        mid = H // 2
        t = time.perf_counter()
        oracle.max_project(vol, iP, iM, box=st["box"], min_val=st["min_val"],
                           max_val=st["max_val"], gamma=st["gamma"], alpha_pow=st["alpha_pow"],
                           max_steps=st["max_steps"], width=W, height=H, rows=(mid, mid + cores),
                           nthreads=cores)
        per_row = (time.perf_counter() - t) / cores
        n = int(max(cores, min(H, per / max(per_row, 1e-6) // cores * cores)))
        r0 = max(0, mid - n // 2)
        r1 = min(H, r0 + n)
        t = time.perf_counter()
        _, a = oracle.max_project(vol, iP, iM, box=st["box"], min_val=st["min_val"],
                                  max_val=st["max_val"], gamma=st["gamma"],
                                  alpha_pow=st["alpha_pow"], max_steps=st["max_steps"], width=W,
                                  height=H, rows=(r0, r1), nthreads=cores)
        dt = time.perf_counter() - t
        hits = int((a[r0:r1] > 0).sum())
        total_s += hits * per_ray
        total_t += dt
        parts.append("%s %dx%d rows %d-%d" % (r.volume, W, H, r0, r1 - 1))
    return {"value": round(total_s / max(total_t, 1e-9) / 1e9, 4), "unit": "Gsamples/s",
            "cores": cores, "kind": "port", "cpu": cpu_model(),
            "sample": "oracle max projection of %s (%.1f s)" % ("; ".join(parts), total_t)}


# ------------------------------------------------------------------------------ learned
def bench_learned(args):
    import torch
    from spimagine_amd import _lib, learned, synth
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices

    shape = (512, 512, 512)
    W = H = 1024
    S = synth.steps_for_shape(shape)
    ctx = _lib.Context(0)
    # held out: the networks are trained on seeds 1.. of the same recipe
    vt = next(synth.blob_series_device(shape, 2000, (3., 10.), (500., 3000.), 100., True, 5., 1,
                                       seed=0))
    u = vt.to(torch.int32) & 0xFFFF
    lo, hi = float(u.min()), float(u.max())
    del u
    slot = ctx.volume_add(dev_ptr=vt.data_ptr(), dtype=np.uint16, shape=shape)
    del vt
    wparams, winputs = learned.load_weights(args.weights)
    lm = learned.LearnedMIP(ctx, params=wparams, inputs=winputs)
    lm.attach(slot)
    P = mat4_perspective(45, 1., .1, 10)
    params = []
    for M in synth.random_poses(8, seed=0):
        iP, iM = render_matrices(M, P, shape[::-1])
        params.append(_lib.make_params(iP, iM, BOX, lo, hi, 1., 0., .1, 21, 30, S,
                                       _lib.FILTER_LINEAR))
    rep = lm.validate(slot, params, W, H)
    stream = torch.cuda.Stream()
    out = torch.empty((H, W), device="cuda")
    alpha = torch.empty_like(out)

    def timed(fn, reps):
        for p in params:  # warm
            fn(p)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            for p in params:
                fn(p)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / (reps * len(params))

    reps = max(1, args.steps // len(params))
    exact_ms = timed(lambda p: ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H,
                                               out.data_ptr(), alpha.data_ptr(),
                                               stream=stream.cuda_stream), reps)
    learned_ms = timed(lambda p: lm.render(slot, p, W, H, out.data_ptr(), stream.cuda_stream), reps)
    frames_buf = torch.empty((len(params), H, W), device="cuda")
    outs = list(frames_buf)  # consecutive views: the network writes them in place
    alphas = [torch.empty_like(o) for o in outs]
    # the 8 learned frames with one network pass (render_many)
    lm.render_many(slot, params, W, H, outs, stream)  # warm
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        lm.render_many(slot, params, W, H, outs, stream)
    e1.record(stream)
    stream.synchronize()
    learned_batch_ms = e0.elapsed_time(e1) / (reps * len(params))
    planes = torch.empty((1, 2, H // 2, W // 2), device="cuda")
    lm.coarse(slot, params[0], W, H, planes, 0, stream.cuda_stream)
    coarse_ms = timed(lambda p: lm.coarse(slot, p, W, H, planes, 0, stream.cuda_stream), reps)
    net_ms = timed(lambda p: lm.net.forward(planes.data_ptr(), 1, H // 2, W // 2, out.data_ptr(),
                                            stream.cuda_stream), reps)
    # per-frame judged rendering of the 8 poses as one batch per call, the calls queued
    # back to back (the judge decides on the device; one host synchronisation at the end):
    # wall clock per frame, the exact re-renders of rejected frames included

    def judged_run(stride, seg, slab_kb=None, concurrent=True, z=learned.JUDGE_Z):
        lm.judge_stride, lm.judge_seg, lm.judge_slab_kb = stride, seg, slab_kb
        lm.judge_concurrent, lm.judge_z = concurrent, z
        lm.render_judged_many(slot, params, outs, alphas, stream)  # warm
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = lm.render_judged_many(slot, params, outs, alphas, stream, wait=False)
        t1 = time.perf_counter()
        stream.synchronize()
        n = reps * len(params)
        log("judged: host enqueue %.4f ms/frame" % (1e3 * (t1 - t0) / n))
        return 1e3 * (time.perf_counter() - t0) / n, res

    # the judge's sampling (one pixel in stride x stride) and how its sparse frames render:
    # split rays (segments of seg samples) or one persistent batch launch (19 KiB slabs
    # or global gathers: the rays run stride pixels apart)
    sweep = []
    for stride, seg, slab, conc, z in () if args.no_sweep else (
                                       (8, 64, None, False, 3.), (8, 32, None, False, 3.),
                                       (8, 64, None, True, 3.), (8, 64, 0, False, 3.),
                                       (16, 64, None, False, 3.), (16, 64, None, False, 2.),
                                       (16, 64, None, True, 3.), (8, None, None, True, 3.),
                                       (16, None, None, True, 3.)):
        ms, res = judged_run(stride, seg, slab, conc, z)
        slab_kb = ctx.get_tuning("slab_kb") if slab is None else slab
        sweep.append({"stride": stride, "z": z, "concurrent": conc,
                      "sparse": "split rays, %d-sample segments, slab_kb %d"
                      % (seg, slab_kb) if seg else "batch launch, slab_kb %d" % slab_kb,
                      "ms_per_frame_wall": round(ms, 4), "learned_kept": res.kept(),
                      "estimated_psnr_db": [round(e, 2) for e in res.estimates()],
                      "bound_psnr_db": [round(e, 2) for e in res.bounds()]})
        log("judge stride %d seg %s slab %s concurrent %s z %g: %.4f ms/frame, kept %s"
            % (stride, seg, slab, conc, z, ms, res.kept()))
    # the reported run: the library's defaults
    judged_ms, res = judged_run(learned.JUDGE_STRIDE, learned.JUDGE_SEG)
    kept, est = res.kept(), res.estimates()
    # the judged frames (learned or fallen back) against the exact ones
    judged_err = []
    for p, o in zip(params, outs):
        ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, out.data_ptr(), alpha.data_ptr(),
                        stream=stream.cuda_stream)
        stream.synchronize()
        e, g = out.cpu().numpy(), o.cpu().numpy()
        judged_err.append((learned.psnr(g, e), float(np.abs(g - e).max())))
    cpu = None
    if not args.no_cpu:
        vol_host = ctx.volume_read(slot)
        rec, _ = cpu_baseline(vol_host, list(params[0].inv_p), list(params[0].inv_m), lo, hi, S,
                              W, H, args.cpu_seconds, np.zeros((H, W), np.float32),
                              np.zeros((H, W), np.float32), "the exact frame of pose 0")
        rows = rec["rows"]
        cpu = dict(rec, value=round(rec["reps"] * (rows[1] - rows[0] + 1) / rec["seconds"] / H,
                                    4), unit="frames/s (exact frames)")
    # matrix-core work of one network pass: 32x32x9 MACs per pixel in 4 layers at h*w,
    # 3 at h*w/4 and 2 at h*w/16, here h = H/2
    px = (H // 2) * (W // 2)
    flops = 2 * 9 * 32 * 32 * px * (4 + 3 / 4. + 2 / 16.)
    achieved = flops / (net_ms * 1e-3) / 1e12
    ok = lm.enabled[slot]
    w = np.load(args.weights)
    print(json.dumps({
        "metric": "learned reconstruction: frames/s vs the exact MIP and PSNR per frame, "
                  "512^3 u16 -> 1024^2", "value": round(1e3 / judged_ms, 2), "unit": "frames/s",
        "n_gpus": 1, "steps": reps * len(params), "warmup": len(params),
        "ms_per_step": round(judged_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "MIP 512^3 uint16 (2000 blobs, Poisson + read noise, held-out seed) "
                               "-> 1024^2, max_steps %d, 8 seeded poses; learned = two coarse "
                               "512^2 planes at max_steps %d (inputs: %s%s) + conv-net"
                               % (S, S // 4, lm.inputs,
                                  ", one pair render of the interleaved 2x mean/max pooled volume"
                                  if lm.attach(slot)[0] == lm.attach(slot)[1] else ""),
                   "training_pairs": int(w["pairs"]) if "pairs" in w else None},
        "exact_ms": round(exact_ms, 4), "exact_frames_per_s": round(1e3 / exact_ms, 2),
        "learned_ms": round(learned_ms, 4), "speedup": round(exact_ms / learned_ms, 3),
        "learned_batch_ms": round(learned_batch_ms, 4),
        "speedup_batch": round(exact_ms / learned_batch_ms, 3),
        "coarse_ms": round(coarse_ms, 4), "net_ms": round(net_ms, 4),
        "psnr_db": [round(r["psnr"], 2) for r in rep],
        "max_abs_err": [round(r["max_abs"], 4) for r in rep],
        "psnr_threshold_db": lm.threshold, "all_validation_frames_pass": ok,
        "judged": {"estimated_psnr_db": [round(e, 2) for e in est],
                   "bound_psnr_db": [round(e, 2) for e in res.bounds()],
                   "acceptance_db": lm.threshold + lm.margin, "z": lm.judge_z,
                   "learned_kept": kept,
                   "false_accepts": int(sum(1 for k, r in zip(kept, rep)
                                            if k and r["psnr"] < lm.threshold)),
                   "psnr_db": [round(v[0], 2) for v in judged_err],
                   "max_abs_err": [round(v[1], 4) for v in judged_err],
                   "ms_per_frame_wall": round(judged_ms, 4),
                   "frames_per_s": round(1e3 / judged_ms, 2),
                   "speedup_vs_exact": round(exact_ms / judged_ms, 3),
                   "stride": learned.JUDGE_STRIDE, "segment": learned.JUDGE_SEG,
                   "sweep": sweep},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(achieved / 2500., 5), "traffic": None,
                     "kernel": "network forward (all launches)"},
        "cpu_baseline": cpu}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("headline", "iso", "batch", "serve", "learned"),
                    default="headline")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--latency-frames", type=int, default=104)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--skip-frames", type=int, default=5,
                    help="frames per pose timed with skipping off and on")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="learned: time the default judge only (no judge-settings sweep)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch tuning passed to spr_set_tuning, e.g. tile_map=1 "
                         "(batch: replaces the autotuner)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run a plain device copy (torch.neg on int16) of this many MiB, "
                         "the reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--timepoints", type=int, default=64)
    ap.add_argument("--poses", type=int, default=16)
    ap.add_argument("--frame", type=int, default=1024)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--qps", type=float, default=60.)
    ap.add_argument("--max-batch", type=int, default=16)
    ap.add_argument("--workers", type=int, default=2)
    ap.add_argument("--hbm-budget-gib", type=float, default=0.)
    ap.add_argument("--weights", default=os.path.join(ROOT, "spimagine_amd", "weights",
                                                      "mip_x2_q4.npz"))
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the ray caster runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    defaults = {"headline": (50, 5), "iso": (40, 8), "batch": (3, 1), "serve": (None, 10),
                "learned": (48, 8)}[args.workload]
    if args.steps is None:
        args.steps = defaults[0]
    if args.warmup is None:
        args.warmup = defaults[1]
    {"headline": bench_headline, "iso": bench_iso, "batch": bench_batch, "serve": bench_serve,
     "learned": bench_learned}[args.workload](args)


if __name__ == "__main__":
    main()
