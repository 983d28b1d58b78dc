"""Headline benchmark: MIP of a 1024x1024x512 uint16 volume into a 1024^2 frame.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]

One step = one max-projection frame (one kernel launch) over the resident volume:
8000 seeded Gaussian blobs, background 100, Poisson + read noise; perspective 45 deg,
modelView translate(0,0,-4).rotY(30 deg), max_steps 1536 (each hitting ray takes
(1536/16+1)*16 = 1552 samples, <= 1 voxel per step), min/max = data min/max, linear.

Prints ONE JSON line.  value = Gsamples/s over the timed steps, where the sample count
is the reference's (rays hitting the box x 1552), independent of the implementation.
roofline.achieved = algorithmic bytes per launch (samples x 8 corners x 2 B + the two
float32 output planes) / average kernel duration (HIP events on the launch stream), over
the 8 TB/s HBM peak.  cpu_baseline = the C oracle (oracle/spimref.c, OpenMP) timed on a
bounded block of rows of the same frame on the host cores.
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
PMC_FILE = os.path.join(ROOT, "profiles", "mip_headline_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--latency-frames", type=int, default=104)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch tuning passed to spr_set_tuning, e.g. tile_map=1")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run a plain device copy (torch.neg on int16) of this many MiB, "
                         "the reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the ray caster runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

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
    box = (-1, 1, -1, 1, -1, 1)

    def set_pose(iP, iM):
        ctx.set_params(iP, iM, box, lo, hi, 1., 0., .1, 21, 30, S, _lib.FILTER_LINEAR)

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
    poses = synth.random_poses(8, seed=0)
    mats = [render_matrices(Mi, P, (nx, ny, nz)) for Mi in poses]
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

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib << 19  # int16 elements
        src = torch.ones(n, dtype=torch.int16, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    b_alg = n_samp * 8 * 2 + W * H * 4 * 2
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PMC_FILE):
        try:
            traffic = json.load(open(PMC_FILE)).get("traffic_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel_ms": round(kernel_ms, 4), "algorithmic_bytes": int(b_alg)}

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        vol = synth.device_volume_to_numpy(vt, np.uint16)
        cores = min(len(os.sched_getaffinity(0)), 16)
        mid = H // 2

        def run_rows(r0, r1):
            t = time.perf_counter()
            o, a = oracle.max_project(vol, invP, invM, min_val=lo, max_val=hi, max_steps=S,
                                      width=W, height=H, rows=(r0, r1), nthreads=cores)
            return time.perf_counter() - t, o, a

        log("bench: cpu baseline (oracle, %d threads)" % cores)
        tcal, o, a = run_rows(mid, mid + cores)
        per_row = tcal / cores
        nrows = int(max(cores, min(H, (args.cpu_seconds / max(per_row, 1e-6)) // cores * cores)))
        r0 = max(0, mid - nrows // 2)
        r1 = min(H, r0 + nrows)
        tcpu, o, a = run_rows(r0, r1)
        hits = int((a[r0:r1] > 0).sum())
        cpu_samples = hits * per_ray
        cpu = {"value": round(cpu_samples / tcpu / 1e9, 4), "unit": "Gsamples/s", "cores": cores,
               "kind": "port",
               "sample": "rows %d-%d of the same 1024^2 frame (%d hit rays, %d samples), %.1f s"
                         % (r0, r1 - 1, hits, cpu_samples, tcpu)}
        parity = float(np.abs(gpu_out[r0:r1] - o[r0:r1]).max())
        if not np.array_equal(gpu_alpha[r0:r1], a[r0:r1]):
            parity = float("inf")

    result = {
        "metric": METRIC,
        "value": round(gsps, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "MIP 1024x1024x512 uint16 volume (8000 seeded blobs, Poisson + "
                               "read noise) -> 1024x1024 float32, max_steps 1536, perspective "
                               "45 deg, translate(0,0,-4).rotY(30 deg), linear sampler",
                   "voxel": "u16", "frame": [W, H], "max_steps": S,
                   "samples_per_frame": int(n_samp), "samples_per_hit_ray": per_ray,
                   "tuning": {k: ctx.get_tuning(k) for k in _lib.TUNING_KEYS}},
        "frames_per_s": round(1e3 / ms_per_step, 2),
        "latency_ms": {"p50": round(pct(lat, 50), 4), "p99": round(pct(lat, 99), 4),
                       "frames": len(lat_ev), "poses": 8,
                       "gsamples_per_s": round(pose_gsps, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_max_abs_diff_vs_oracle": parity,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
