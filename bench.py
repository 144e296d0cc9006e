#!/usr/bin/env python3
"""Benchmark: 8-lambda path segments per second of the HIP transport path on one MI355X.

Workload (single GPU job): a uniaxial, pleochroic sapphire (Sellmeier n_o ~ 1.768,
n_e = n_o - 0.0081, seeded optic axis, 8 Gaussian absorption bands per axis with peaks
0.01-2 cm^-1) cut as a 69-plane oval brilliant, lit by a 2048x1024 procedural HDR
lat-long environment, rendered at 1920x1080 with 256 samples per pixel and at most 32
bounces.  One step = one full-frame render (530.8 M paths) accumulated into a
device-resident film; inputs are uploaded before the timed region.

Prints one JSON line (rank 0).  ``value`` is whole-job segments per second over the K
timed steps.  ``roofline`` reports the trace kernel's algorithmic FP32 rate (exact
per-class event counters x the op-count model of gemray_studio_amd/op_counts.py) over its
average launch time from HIP events on the library's stream.  ``cpu_baseline`` times the
oracle (the CPU restatement of the reference tracer) on a bounded window of the same
frame, on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "8-λ path segments/sec on 1 MI355X; p99 job latency, batched vs plain"

from gemray_studio_amd.workloads import (batched_jobs, cpu_config_scene, served_jobs, single_job_scene,  # noqa: E402
                                         stress_scene)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def run_plain(gpu, jobs):
    """The plain per-job path: one gr_accumulate per job, in order, into a host film
    (scene and environment uploaded per call, film read back per call)."""
    service, segs = [], 0
    t0 = time.perf_counter()
    for sc, first, spp in jobs:
        film = np.zeros((sc.width * sc.height, 3), np.float32)
        t = time.perf_counter()
        gpu.accumulate(sc, first, spp, film)
        service.append(time.perf_counter() - t)
        segs += gpu.last_stats()["segments"]
    return time.perf_counter() - t0, service, segs


def run_batched(gpu, jobs, arrivals=None):
    """Submits every job (at its arrival time, if given) to the job loop and waits for all;
    returns wall seconds from the first submission, per-job latencies and segments."""
    from gemray_studio_amd import _lib

    gpu.metrics_reset()
    ids = []
    t0 = time.perf_counter()
    for k, (sc, first, spp) in enumerate(jobs):
        if arrivals is not None:
            delay = t0 + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        ids.append(gpu.submit(sc, first, spp))
    for jid in ids:
        if gpu.wait(jid, 600000) != _lib.GR_OK:
            raise RuntimeError(f"job {jid} did not complete")
    wall = time.perf_counter() - t0
    infos = [gpu.job_info(jid) for jid in ids]
    for jid in ids:
        gpu.release(jid)
    m = gpu.metrics()
    return wall, [i["latency_ms"] / 1e3 for i in infos], sum(i["events"]["segment"] for i in infos), m


def fifo_latencies(arrivals, service):
    """Submit-to-completion latencies of a single FIFO server with these service times."""
    free_at, out = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free_at)
        free_at = start + s
        out.append(free_at - a)
    return out


def pct(v, q):
    v = sorted(v)
    return v[min(len(v), max(1, int(np.ceil(q * len(v))))) - 1]


def cpu_baseline(scene, cam_args, spp, budget_s):
    """Oracle on row windows of the same frame (spp samples per pass), grown until the run
    takes about budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from gemray_studio_amd import materials as M

    threads = len(os.sched_getaffinity(0))
    cam = po.camera_new(*cam_args)
    env = po.map_env(scene.environment.rgb)
    mat = M.to_struct(scene.material)
    w, h = scene.width, scene.height
    rows = 2
    total_segs = 0
    total_paths = 0
    total_rows = 0
    elapsed = 0.0
    y0 = h // 2
    sample_offset = 0
    # Row windows from the frame centre downwards, wrapping to the top; a full pass moves
    # on to the next 16 samples, so every pixel is weighted like the GPU frame.
    while elapsed < budget_s:
        n = min(rows, (h if y0 >= h // 2 else h // 2) - y0)
        p0, p1 = y0 * w, (y0 + n) * w
        t0 = time.perf_counter()
        _, _, _, segs = po.render(cam, w, h, scene.planes, mat, scene.max_bounces, env,
                                  sample_offset, spp, pixel_begin=p0, pixel_end=p1, threads=threads)
        elapsed += time.perf_counter() - t0
        total_segs += segs
        total_paths += (p1 - p0) * spp
        total_rows += n
        y0 += n
        if y0 == h:
            y0 = 0
        if y0 == h // 2:
            sample_offset += spp
        rows = min(rows * 2, 64)
    return {
        "value": total_segs / elapsed,
        "unit": "segments/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (C++ restatement of trace_spectral_ray) on {total_rows} rows of the "
                  f"{w}x{h} frame (from row {h // 2}, wrapping) at {spp} spp per row pass: "
                  f"{total_paths} paths, {total_segs} segments in {elapsed:.1f} s on {threads} threads",
        "paths_per_s": total_paths / elapsed,
    }


def metrics_workload(gpu, args):
    """Gemological metrics of the batched workload's 64 stones on the device: each stone at
    its job's camera pose under its studio lights, plus its 19-pitch angular profile at that
    azimuth (evaluate_gem_optical_metrics and evaluate_angular_profile_at_azimuth,
    color/metrics.rs) -- 20 poses per stone, one gr_optical_metrics call per stone.  The
    oracle (the C++ restatement, one thread) times a bounded sample of the same calls."""
    from gemray_studio_amd import backend as B

    poses = []
    jobs = batched_jobs(n_jobs=args.batched_jobs, poses=poses)
    calls = []
    for (sc, _, _), (yaw, pitch, ly, lp) in zip(jobs, poses):
        calls.append((sc.planes, sc.material, [(yaw, pitch, ly, lp)] +
                      [(yaw, float(p), ly, lp) for p in B.PROFILE_PITCHES_RAD]))
    gpu.optical_metrics_batch(calls[:2])  # first-call setup outside the timed region
    t0 = time.perf_counter()
    per_stone = [gpu.optical_metrics(*c) for c in calls]
    per_stone_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    results = gpu.optical_metrics_batch(calls)
    gpu_s = time.perf_counter() - t0
    if results != per_stone:
        raise RuntimeError("metrics: one launch for all stones differs from one call per stone")
    n_eval = sum(len(c[2]) for c in calls)
    out = {
        "workload": f"{len(calls)} batched-workload stones (17-181 planes) x (job pose + 19-pitch angular "
                    "profile) = 20 poses per stone, all stones in one device launch (gr_optical_metrics_batch)",
        "evaluations": n_eval, "gpu_s": gpu_s, "evaluations_per_s": n_eval / gpu_s,
        "per_stone_calls_evaluations_per_s": n_eval / per_stone_s,
        "fire_index_mean": float(np.mean([r["fire_index"] for rs in results for r in rs])),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        from gemray_studio_amd import materials as M

        t0 = time.perf_counter()
        n_cpu, identical = 0, True
        for (planes, material, ps), rs in zip(calls, results):
            m = M.to_struct(material)
            for q, r in zip(ps, rs):
                v = po.gem_optical_metrics(planes, m, *q)
                identical &= all(np.float32(r[n]) == v[i] for i, n in enumerate(po.METRIC_NAMES))
                n_cpu += 1
            if time.perf_counter() - t0 > 3.0:
                break
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"evaluations_per_s": n_cpu / cpu_s, "cores": 1, "kind": "port",
                               "sample": f"oracle evaluate_gem_optical_metrics, first {n_cpu} of the same "
                                         f"evaluations in {cpu_s:.1f} s on one thread"}
        out["identical_to_oracle_on_sample"] = bool(identical)
    return out


def feature_workloads(gpu, scene, args):
    """The material branches and passes beside the single job, each on the single-job frame
    (oval brilliant, 2048x1024 HDR map, 1920x1080, depth 32) with one thing changed, timed as
    one 16-spp launch after a warm-up: biaxial stones (Alexandrite, trichroic; Tanzanite),
    Henyey-Greenstein scattering (Ruby with sigma_s 0.8, g 0.4), frosted near-vertical
    (girdle) facets on the single-job sapphire; and the denoiser's guide pass over the frame
    (gr_guide_buffers, host copies of its 41 MB of outputs included)."""
    import dataclasses
    from gemray_studio_amd import materials as M

    spp = 16
    girdle = (np.abs(scene.planes[:, 1]) < 0.2).astype(np.uint32)
    variants = [
        ("biaxial_alexandrite", dataclasses.replace(scene, material=M.by_name("Alexandrite"))),
        ("biaxial_tanzanite", dataclasses.replace(scene, material=M.by_name("Tanzanite"))),
        ("scattering_ruby", dataclasses.replace(scene, material=dataclasses.replace(
            M.by_name("Ruby"), scattering_sigma_s=0.8, scattering_g=0.4))),
        ("frosted_girdle_sapphire", dataclasses.replace(scene, facet_finishes=girdle)),
    ]
    out = {"workload": "single-job frame (1920x1080, depth 32) with the material or finishes changed, "
                       f"one {spp}-spp launch each; guide pass over the same frame"}
    film = gpu.film(scene.width, scene.height)
    for name, sc in variants:
        ds = gpu.upload(sc)
        gpu.render(ds, 0, 2, film)
        best = None
        for k in range(2):
            gpu.render(ds, 2 + k * spp, spp, film)
            st = gpu.last_stats()
            if best is None or st["total_ms"] < best["total_ms"]:
                best = st
        out[name] = {"segments_per_s": best["segments"] / (best["total_ms"] * 1e-3),
                     "segments_per_path": best["segments"] / (scene.width * scene.height * spp),
                     "launch_ms": best["total_ms"]}
        print(f"bench: {name} {out[name]['segments_per_s'] / 1e9:.2f} G segments/s", file=sys.stderr, flush=True)
    gpu.guide_buffers(scene.camera, scene.width, scene.height, scene.planes)
    t0 = time.perf_counter()
    for _ in range(5):
        gpu.guide_buffers(scene.camera, scene.width, scene.height, scene.planes)
    gs = (time.perf_counter() - t0) / 5
    out["guide_pass"] = {"pixels_per_s": scene.width * scene.height / gs, "call_ms": gs * 1e3}
    return out


def job_workloads(gpu, env_map, args):
    """Batched (all jobs submitted at once) and served (Poisson arrivals) throughput and
    latency of the job loop, each beside the plain per-job path on the same jobs."""
    jobs = batched_jobs(n_jobs=args.batched_jobs)
    paths = sum(sc.width * sc.height * spp for sc, _, spp in jobs)
    print(f"bench: batched workload, {len(jobs)} jobs, {paths} paths", file=sys.stderr, flush=True)
    plain_s, plain_service, plain_segs = run_plain(gpu, jobs)
    print(f"bench: plain per-job path {plain_s:.2f} s", file=sys.stderr, flush=True)
    b_wall, b_lat, b_segs, b_m = run_batched(gpu, jobs)
    print(f"bench: job loop {b_wall:.2f} s", file=sys.stderr, flush=True)
    if b_segs != plain_segs:
        raise RuntimeError("batched and plain paths traced different work")
    batched = {
        "workload": f"{len(jobs)} seeded convex stones (17-181 planes, +-3 deg), isotropic/uniaxial mix, "
                    "128^2-1024^2 px, 64-1024 spp (log-uniform), studio presets, depth 16, all submitted at t=0",
        "jobs": len(jobs), "paths": paths, "segments": b_segs,
        "plain_s": plain_s, "plain_segments_per_s": plain_segs / plain_s,
        "batched_s": b_wall, "batched_segments_per_s": b_segs / b_wall,
        "speedup": plain_s / b_wall,
        "batched_busy_s": b_m["busy_ms"] / 1e3, "ticks": b_m["ticks"],
        "plain_latency_p50_ms": 1e3 * pct(np.cumsum(plain_service), 0.5),
        "plain_latency_p99_ms": 1e3 * pct(np.cumsum(plain_service), 0.99),
        "batched_latency_p50_ms": b_m["latency_p50_ms"], "batched_latency_p99_ms": b_m["latency_p99_ms"],
    }
    sjobs = served_jobs(env_map, n_jobs=args.served_jobs)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / args.arrival_rate, size=len(sjobs)))
    arrivals -= arrivals[0]
    s_plain_s, s_service, s_plain_segs = run_plain(gpu, sjobs)
    plain_lat = fifo_latencies(arrivals, s_service)
    print(f"bench: served workload, plain per-job path {s_plain_s:.2f} s", file=sys.stderr, flush=True)
    s_wall, s_lat, s_segs, s_m = run_batched(gpu, sjobs, arrivals=arrivals)
    print(f"bench: served workload, job loop {s_wall:.2f} s", file=sys.stderr, flush=True)
    served = {
        "workload": f"{len(sjobs)} jobs, Poisson arrivals at {args.arrival_rate:g}/s; batched-workload "
                    "per-job distribution with the single-job material/env family (seeded pleochroic "
                    "sapphire, 2048x1024 HDR map, depth 32), 10% at 1920x1080 x 512 spp",
        "jobs": len(sjobs), "arrival_rate_per_s": args.arrival_rate,
        "offered_load": float(np.sum(s_service) / (arrivals[-1] + 1.0 / args.arrival_rate)),
        "segments": s_segs,
        "served_segments_per_s": s_segs / s_wall, "served_wall_s": s_wall,
        "latency_p50_ms": 1e3 * pct(s_lat, 0.5), "latency_p99_ms": 1e3 * pct(s_lat, 0.99),
        "plain_fifo_latency_p50_ms": 1e3 * pct(plain_lat, 0.5),
        "plain_fifo_latency_p99_ms": 1e3 * pct(plain_lat, 0.99),
        "plain_segments_per_s": s_plain_segs / s_plain_s,
        "plain_fifo_wall_s": float(arrivals[-1] + plain_lat[-1]),
    }
    previews = preview_workload(gpu, env_map, args)
    return batched, served, previews


def preview_workload(gpu, env_map, args):
    """Interactive previews (workloads.preview_jobs): many small frames at few samples that
    share one device-resident map, submitted all at once (batched) and as a Poisson stream at
    3x the plain path's service rate (served), beside the plain per-job path (gr_accumulate
    per job, the same device-resident map)."""
    from gemray_studio_amd import backend as B
    from gemray_studio_amd.workloads import preview_jobs

    dev = gpu.upload_env_map(env_map)
    try:
        jobs = preview_jobs(B.HdrMap(env_map, device=dev), n_jobs=args.preview_jobs)
        paths = sum(sc.width * sc.height * spp for sc, _, spp in jobs)
        run_plain(gpu, jobs[:8])  # first-call setup outside the measurement
        plain_s, service, plain_segs = run_plain(gpu, jobs)
        print(f"bench: previews, {len(jobs)} jobs, plain per-job path {plain_s:.2f} s", file=sys.stderr, flush=True)
        # The loop's stream-ordered memory pool and pinned staging grow to the size of 400
        # jobs admitted at once on first use; a serving loop holds them warm, as the plain
        # path's first calls above are kept out of its measurement.
        run_batched(gpu, jobs)
        b_wall, b_lat, b_segs, b_m = run_batched(gpu, jobs)
        if b_segs != plain_segs:
            raise RuntimeError("preview jobs: batched and plain paths traced different work")
        rate = 3.0 * len(jobs) / plain_s
        rng = np.random.default_rng(7)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(jobs)))
        arrivals -= arrivals[0]
        s_wall, s_lat, s_segs, s_m = run_batched(gpu, jobs, arrivals=arrivals)
        plain_lat = fifo_latencies(arrivals, service)
        print(f"bench: previews, job loop {b_wall:.2f} s batched, {s_wall:.2f} s served", file=sys.stderr,
              flush=True)
    finally:
        dev.release()
    return {
        "workload": f"{len(jobs)} preview jobs: 96^2-256^2 px, 4-32 spp (log-uniform), seeded convex stones "
                    "(17-181 planes), seeded pleochroic sapphires, one device-resident 2048x1024 HDR map, depth 32",
        "jobs": len(jobs), "paths": paths, "segments": b_segs,
        "plain_s": plain_s, "plain_segments_per_s": plain_segs / plain_s,
        "plain_jobs_per_s": len(jobs) / plain_s,
        "batched_s": b_wall, "batched_segments_per_s": b_segs / b_wall, "batched_speedup": plain_s / b_wall,
        "batched_latency_p99_ms": b_m["latency_p99_ms"],
        "served_arrival_rate_per_s": rate, "served_wall_s": s_wall,
        "served_segments_per_s": s_segs / s_wall,
        "served_latency_p50_ms": 1e3 * pct(s_lat, 0.5), "served_latency_p99_ms": 1e3 * pct(s_lat, 0.99),
        "plain_fifo_latency_p50_ms": 1e3 * pct(plain_lat, 0.5),
        "plain_fifo_latency_p99_ms": 1e3 * pct(plain_lat, 0.99),
        "plain_fifo_wall_s": float(arrivals[-1] + plain_lat[-1]),
        "served_throughput_over_plain": (s_segs / s_wall) / (plain_segs / float(arrivals[-1] + plain_lat[-1])),
    }


def config_workload(gpu, scene, cam_args, spp, steps, cpu_budget_s, cpu_spp, no_cpu, label, autotune=True):
    """One more configuration, timed like the single job (device-resident scene and film,
    steps of spp samples per pixel, the autotuner's choice for it picked before the timed
    region), with the oracle on a bounded window of the same frame (row windows from the
    frame centre, cpu_spp samples, about cpu_budget_s seconds)."""
    dscene = gpu.upload(scene)
    film = gpu.film(scene.width, scene.height)
    tuned = None
    try:
        if autotune:
            tuned = gpu.autotune(dscene, spp)
            gpu.set_autotune(True)
        else:
            gpu.render(dscene, 0, min(spp, 4), film)  # first-launch costs (the probes pay them)
        segs = paths = 0
        ms = 0.0
        t0 = time.perf_counter()
        for k in range(steps):
            gpu.render(dscene, (k + 1) * spp, spp, film)
            st = gpu.last_stats()
            segs += st["segments"]
            paths += st["paths"]
            ms += st["total_ms"]
        wall = time.perf_counter() - t0
    finally:
        gpu.set_autotune(False)
        film.free()
        dscene.free()
    out = {"workload": label, "width": scene.width, "height": scene.height, "spp_per_step": spp, "steps": steps,
           "autotuned": tuned,
           "n_planes": int(scene.planes.shape[0]), "max_bounces": scene.max_bounces,
           "segments_per_s": segs / wall, "paths_per_s": paths / wall, "segments_per_path": segs / max(paths, 1),
           "ms_per_step": wall * 1e3 / steps, "device_ms_per_step": ms / steps}
    print(f"bench: {label}: {out['segments_per_s'] / 1e9:.2f} G segments/s", file=sys.stderr, flush=True)
    if not no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cam_args, cpu_spp, cpu_budget_s)
        out["gpu_over_cpu"] = out["segments_per_s"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256, help="samples per pixel per step")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--trace-mode", choices=["plain", "wavefront"], default="plain")
    ap.add_argument("--no-jobs", action="store_true", help="skip the batched and served workloads")
    ap.add_argument("--no-autotune", action="store_true", help="time the context's default configuration")
    ap.add_argument("--batched-jobs", type=int, default=64)
    ap.add_argument("--served-jobs", type=int, default=40)
    ap.add_argument("--arrival-rate", type=float, default=20.0, help="served jobs per second (Poisson)")
    ap.add_argument("--preview-jobs", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a run", file=sys.stderr)
        sys.exit(2)

    from gemray_studio_amd import backend as B
    from gemray_studio_amd import op_counts

    scene, cam_args = single_job_scene(args.width, args.height)
    gpu = B.GpuBackend.acquire()
    if not gpu.available:
        print("bench.py: no gfx950 device", file=sys.stderr)
        sys.exit(1)
    print(f"bench: {gpu.adapter_label()}; {args.width}x{args.height} x {args.spp} spp/step, "
          f"P={scene.planes.shape[0]}", file=sys.stderr, flush=True)
    gpu.set_trace_mode(args.trace_mode)
    dscene = gpu.upload(scene)
    film = gpu.film(scene.width, scene.height)
    tuned = None
    if not args.no_autotune:
        # The autotuner's choice for this scene class and size, picked from probe renders
        # before the timed region; the timed steps then use it (gr_set_autotune).
        tuned = gpu.autotune(dscene, args.spp)
        gpu.set_autotune(True)
        print(f"bench: autotuned {tuned}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        gpu.render(dscene, i * args.spp, args.spp, film)
        print(f"bench: warmup {i} {gpu.last_stats()['total_ms']:.0f} ms", file=sys.stderr, flush=True)
    segments = 0
    paths = 0
    trace_ms = 0.0
    finish_ms = 0.0
    device_ms = 0.0
    launches = 0
    events = None
    t0 = time.perf_counter()
    for k in range(args.steps):
        gpu.render(dscene, (args.warmup + k) * args.spp, args.spp, film)
        st = gpu.last_stats()
        segments += st["segments"]
        paths += st["paths"]
        trace_ms += st["trace_kernel_ms"]
        finish_ms += st["finish_kernel_ms"]
        device_ms += st["total_ms"]
        tables = (st["spectral_table"], st["material_table"])
        launches += st["launches"]
        if events is None:
            events = dict(st["events"])
        else:
            for n, v in st["events"].items():
                events[n] += v
        print(f"bench: step {k} {st['total_ms']:.0f} ms, {st['segments']} segments", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    film.free()
    dscene.free()
    gpu.set_autotune(False)

    m = scene.material
    model = (events, scene.planes.shape[0], paths, m.is_anisotropic, bool(m.o_bands or m.e_bands), False)
    n_bands = len(m.o_bands) + len(m.e_bands)
    flops = op_counts.algorithmic_flops(*model, n_bands=n_bands, spectral_table=tables[0],
                                        material_table=tables[1])
    flops_uniform = op_counts.algorithmic_flops_uniform(*model, n_bands=n_bands)
    background = events["first_miss"]
    # One launch = one chunk's trace kernel plus its continuation levels (plain mode) or its
    # finish kernel (wavefront mode); the op-count model covers them all.  Consecutive chunks
    # alternate between two streams: a chunk's trace kernel starts when the previous chunk's
    # has ended, and the previous chunk's continuation levels run beside it, so a launch's
    # duration is the device time of the step (HIP events on the library's stream around all
    # of its launches) per launch.
    avg_launch_s = device_ms / 1e3 / launches
    achieved_tf = flops / launches / avg_launch_s / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"{'trace_kernel' if args.trace_mode == 'plain' else 'trace_wave_kernel'}"
                                           "_single_job_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": segments / elapsed,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single job: pleochroic uniaxial sapphire, 69-plane oval brilliant, 2048x1024 "
                        f"procedural HDR env, {args.width}x{args.height} x {args.spp} spp per step, depth 32, "
                        f"framed with the reference's test camera {tuple(cam_args)} (yaw, pitch, distance, fov)",
            "camera": list(cam_args),
            "width": args.width, "height": args.height, "spp_per_step": args.spp,
            "trace_mode": args.trace_mode,
            "autotuned": tuned,
            "max_bounces": scene.max_bounces, "n_planes": int(scene.planes.shape[0]),
            "paths": paths, "segments": segments,
            "segments_per_path": segments / max(paths, 1),
            "paths_per_s": paths / elapsed,
            # Paths whose camera ray misses the stone (one segment each, device-counted), and
            # the throughput of the segments that meet it.
            "background_path_fraction": background / max(paths, 1),
            "stone_segments_per_s": (segments - background) / elapsed,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tf,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tf / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel": ("trace_kernel + its continuation launches (cont_kernel)" if args.trace_mode == "plain"
                       else "trace_wave_kernel + finish_kernel"),
            "avg_launch_ms": avg_launch_s * 1e3,
            # Per chunk, the trace kernel's start to its last continuation level's end (these
            # spans overlap the next chunk's trace kernel, so they sum to more than the step).
            "avg_chunk_span_ms": trace_ms / launches,
            "avg_finish_kernel_ms": finish_ms / launches,
            "launches": launches,
            "algorithmic_flops_per_launch": flops / launches,
            # Hero tables the timed launches read (their lookups are charged no operations).
            "spectral_table": tables[0], "material_table": tables[1],
            # The device's exact per-class event counters, per launch (the model's inputs).
            "events_per_launch": {k: v / launches for k, v in events.items()},
            # The same events charged as if no shortcut ran (every segment the full clip,
            # every path the full setup and finish): an upper bound on the credited work.
            "frac_if_every_segment_ran_the_full_clip": flops_uniform / launches / avg_launch_s / 1e12
            / PEAK_FP32_TFLOPS,
        },
    }
    if not args.no_jobs:
        out["batched"], out["served"], out["previews"] = job_workloads(gpu, scene.environment.rgb, args)
        out["metrics"] = metrics_workload(gpu, args)
        print(f"bench: metrics {out['metrics']['evaluations_per_s']:.0f} evaluations/s", file=sys.stderr,
              flush=True)
        out["features"] = feature_workloads(gpu, scene, args)
        c1, c1_cam = cpu_config_scene()
        out["cpu_configuration"] = config_workload(
            gpu, c1, c1_cam, 16, 5, args.cpu_budget / 2, 16, args.no_cpu,
            "the CPU tracer's configuration: Diamond on the parametric round brilliant (57 facets + 16 girdle "
            "planes), uniform furnace environment, 256x256 x 16 spp, depth 16", not args.no_autotune)
        c5, c5_cam = stress_scene()
        out["stress"] = config_workload(
            gpu, c5, c5_cam, 16, 1, args.cpu_budget / 2, 4, args.no_cpu,
            "stress: Abbe-20, n_e - n_o = 0.17 stone on the 233-plane step cut, 2048x1024 HDR map, 4096x4096 "
            "x 16 spp per step, depth 64", not args.no_autotune)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cam_args, 16, args.cpu_budget)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
