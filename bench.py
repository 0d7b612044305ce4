#!/usr/bin/env python3
"""Benchmark of the path-tracing core on one MI355X.

Workload (default): the ~262k-triangle procedural field scene (512x512-class
noise heightfield sized to 261,920 triangles with 64 icospheres, 256 emissive
triangles), 1920x1080, 8 samples per pixel, max bounces 6. One step = one
render job = every sample of the frame, traced from sample 1 (the scene,
its BVH and light tables are resident in HBM before timing starts).

Prints one JSON line: path samples/s (value), rays/s, p50/p99 job latency,
the roofline of the traversal (extend) kernel measured live with HIP events
on its stream, and the CPU oracle's throughput on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload field|cornell|large|batch|serve|interactive]
                    [--schedule wavefront|mega] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "path-tracer_amd"))

HBM_PEAK_GBS = 8000.0
PAIR_BYTES, TRI_BYTES = 64, 36  # a child pair of 32-B nodes; one 36-B triangle

WORKLOADS = {
    # name: (scene kind, width, height, spp, max_bounces, description)
    "field": ("SCENE_FIELD", 1920, 1080, 8, 6,
              "field-262k: 261,920-triangle noise heightfield + 64 icospheres, 256 emissive tris, 1920x1080, 8 spp, depth 6"),
    "cornell": ("SCENE_CORNELL", 256, 256, 16, 8, "cornell-36: Cornell box, 256x256, 16 spp, depth 8"),
    "large": ("SCENE_LARGE", 3840, 2160, 64, 8,
              "large-4.2M: 4,195,522-triangle heightfield + 4096 icospheres, 8 Perlin textures, 3840x2160, 64 spp, depth 8"),
}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(world, w, h, bounces, samples, repeats, seconds_per_run=4.0):
    """The CPU oracle (C++ restatement of the reference method, OpenMP over
    pixels) on the host cores, on a bounded sample of the benchmark's job: the
    whole frame when one run takes at most about `seconds_per_run` (the 1080p
    job), else a central band of full-width rows sized to that time from a
    short probe, at the job's spp and depth; `repeats` runs, the median is
    reported."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import binding as oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    probe = min(h, max(1, threads))
    t0 = time.perf_counter()
    oracle.render_rows_f32(world.desc, w, h, bounces, samples, (h - probe) // 2, (h - probe) // 2 + probe, threads=threads)
    per_row = (time.perf_counter() - t0) / probe
    rows = min(h, max(probe, int(seconds_per_run / max(per_row, 1e-9))))
    row0 = (h - rows) // 2
    runs = []
    for k in range(max(1, repeats)):
        t0 = time.perf_counter()
        _, st = oracle.render_rows_f32(world.desc, w, h, bounces, samples, row0, row0 + rows, threads=threads)
        runs.append((time.perf_counter() - t0, st))
        print(f"[bench] CPU baseline run {k + 1}: {runs[-1][0]:.2f} s", file=sys.stderr, flush=True)
    runs.sort(key=lambda r: r[0])
    dt, st = runs[len(runs) // 2]
    total = sum(r[0] for r in runs)
    what = (f"the benchmark job ({w}x{h}" if rows == h else
            f"a central band of the benchmark job's frame (rows {row0}-{row0 + rows - 1} of {w}x{h}")
    return {"value": w * rows * samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"{what}, {samples} spp, depth {bounces}) run {len(runs)} times, "
                      f"median {dt:.2f} s of {total:.1f} s total; rays/s "
                      f"{((st['closest_rays'] + st['shadow_rays']) / dt):.4g}"}


def load_traffic(name):
    """HBM bytes per extend launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", f"extend_traffic_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def build_worlds(specs):
    """Prepare many procedural worlds in parallel host threads (the C++ build releases the GIL)."""
    import ptcore
    from concurrent.futures import ThreadPoolExecutor

    def one(spec):
        kind, seed, w, h, tris = spec
        return ptcore.World.generate(kind, seed, w, h, tris)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        return list(ex.map(one, specs))


def mixed_jobs(seed, count):
    """The mixed-scene batch: triangle counts log-uniform over 1k-500k,
    resolutions uniform over {256, 512, 768, 1024}^2, spp uniform over {4, 16, 64}, depth 6."""
    import random

    rng = random.Random(seed)
    jobs = []
    for i in range(count):
        tris = int(math.exp(rng.uniform(math.log(1000), math.log(500000))))
        res = rng.choice([256, 512, 768, 1024])
        spp = rng.choice([4, 16, 64])
        jobs.append(dict(tris=tris, width=res, height=res, spp=spp, bounces=6, seed=1000 + i))
    return jobs


def bench_batch(args):
    """Mixed-scene batch: every job traced by the plain single-scene megakernel one after
    another, then all jobs as one autotuned mixed wavefront batch."""
    import ptcore

    jobs = mixed_jobs(4, args.batch_scenes)
    t0 = time.perf_counter()
    worlds = build_worlds([(ptcore.SCENE_SIZED, j["seed"], j["width"], j["height"], j["tris"]) for j in jobs])
    prep_s = time.perf_counter() - t0
    ctx = ptcore.Context(0)
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    print(f"[bench] batch of {len(jobs)} scenes, {sum(w.desc.triangle_count for w in worlds)} triangles, "
          f"{samples / 1e6:.1f}M path samples, host prep {prep_s:.1f}s", file=sys.stderr, flush=True)

    def tracers(schedule):
        ts = []
        for j, s in zip(jobs, scenes):
            t = ptcore.Tracer(ctx, s, j["width"], j["height"], schedule)
            t.set_max_bounces(j["bounces"])
            ts.append(t)
        return ts

    # plain method: megakernel, one scene at a time
    mk = tracers(ptcore.SCHEDULE_MEGAKERNEL)
    for t, j in zip(mk[:2], jobs[:2]):  # warm the code objects
        t.trace_all(1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for t, j in zip(mk, jobs):
        t.set_sample_count(0)
        t.trace_all(j["spp"])
    ctx.synchronize()
    mega_s = time.perf_counter() - t0
    for t in mk:
        t.close()
    # wavefront, one scene at a time (no batching)
    wf = tracers(ptcore.SCHEDULE_WAVEFRONT)
    ctx.synchronize()
    t0 = time.perf_counter()
    for t, j in zip(wf, jobs):
        t.set_sample_count(0)
        t.trace_all(j["spp"])
    ctx.synchronize()
    single_wf_s = time.perf_counter() - t0
    # batched: autotune, then timed steps of the whole batch
    spp = [j["spp"] for j in jobs]
    tune = ptcore.autotune_batch(wf, spp) if not args.no_autotune else None
    print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
    times = []
    for k in range(args.warmup + args.steps):
        for t in wf:
            t.set_sample_count(0)
        ctx.synchronize()
        s0 = time.perf_counter()
        ptcore.trace_batch(wf, spp)
        ctx.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - s0)
    batch_s = sum(times) / len(times)
    for t in wf:
        t.close()
    value = samples / batch_s
    return {
        "metric": "path samples/s @1080p depth 6; rays/s; p50/p99 render-job latency",
        "value": value, "unit": "path samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": batch_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": f"mixed batch: {len(jobs)} scenes, triangles log-uniform 1k-500k, resolutions "
                               f"{{256..1024}}^2, spp {{4,16,64}}, depth 6",
                   "path_samples_per_step": samples, "wave_config": ptcore.get_wave_config(ctx), "autotune": tune},
        "megakernel_sequential": {"value": samples / mega_s, "seconds": mega_s},
        "wavefront_sequential": {"value": samples / single_wf_s, "seconds": single_wf_s},
        "batched_vs_megakernel": (samples / batch_s) / (samples / mega_s),
    }


def bench_serve(args):
    """Serving: Poisson arrivals of render requests over resident scenes (one
    4.2M-triangle scene and two 262k-triangle scenes); each request draws its
    resolution from {256..1024}^2 and spp from {4, 16, 64}, depth 6. For each
    arrival rate: p50/p99 submit-to-result latency and sustained samples/s."""
    import random
    import threading

    import ptcore

    t0 = time.perf_counter()
    worlds = build_worlds([(ptcore.SCENE_LARGE, 3, 1024, 1024, 0), (ptcore.SCENE_FIELD, 2, 1024, 1024, 0),
                           (ptcore.SCENE_FIELD, 12, 1024, 1024, 0)])
    print(f"[bench] serving pool prepared in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = ptcore.Context(0)
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    results = []
    for lam in args.rates:
        srv = ptcore.Server(ctx, hbm_budget_bytes=200 << 30, max_active_jobs=64)
        rng = random.Random(int(lam * 1000) + 5)
        reqs = [(rng.randrange(len(scenes)), rng.choice([256, 512, 768, 1024]), rng.choice([4, 16, 64]))
                for _ in range(args.jobs_per_rate)]
        gaps = [rng.expovariate(lam) for _ in reqs]
        ids = []

        def producer():
            t_next = time.perf_counter()
            for (si, res, spp), gap in zip(reqs, gaps):
                t_next += gap
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append((srv.submit(scenes[si], res, res, spp, 6), res, spp))

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        lat, first, last, samples = [], None, 0.0, 0
        for jid, res, spp in ids:
            srv.wait(jid, 600_000)
            info = srv.info(jid)
            srv.result(jid)
            lat.append((info["done_s"] - info["submit_s"]) * 1e3)
            first = info["submit_s"] if first is None else min(first, info["submit_s"])
            last = max(last, info["done_s"])
            samples += res * res * spp
        st = srv.stats()
        srv.close()
        r = {"rate_jobs_per_s": lam, "jobs": len(ids), "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
             "samples_per_s": samples / (last - first), "rounds": st["rounds"], "busy_s": st["busy_s"]}
        print(f"[bench] serve {r}", file=sys.stderr, flush=True)
        results.append(r)
    return {"metric": "p50/p99 render-job latency under Poisson arrivals; path samples/s",
            "unit": "ms", "n_gpus": 1, "data": "synthetic (seeded procedural scenes)",
            "config": {"workload": "serving: pool of one 4.2M-tri and two 262k-tri resident scenes; requests "
                                   "{256..1024}^2, spp {4,16,64}, depth 6; 200 GB HBM budget, 64 slots"},
            "sweep": results}


def wave_knobs(args):
    """The wavefront configuration fields set on the command line."""
    return {k: v for k, v in (("traversal", args.traversal), ("sort_key", args.sort), ("lds_stack_depth", args.lds),
                              ("streams", args.streams), ("chunk_paths", args.chunk))
            if v is not None}


def bench_interactive(args):
    """Progressive frames as the reference's interactive mode renders them
    (interactive.rs:554, 611): each frame traces one more sample of every
    pixel (PathTracer::trace with n = 1) and tone-maps the running image for
    display. The RGBA8 frame is copied to host memory here; a window's swap
    chain would take it instead. One step is one frame; both schedules run."""
    import numpy as np
    import ptcore

    W, H, BOUNCES = 1920, 1080, 6
    frames = args.steps if args.steps is not None else 200
    warm = args.warmup if args.warmup is not None else 20
    world = ptcore.World.generate(ptcore.SCENE_FIELD, 2, W, H)
    ctx = ptcore.Context(0)
    scene = ptcore.Scene(ctx, world.desc)
    knobs = wave_knobs(args)
    if knobs:
        ptcore.set_wave_config(ctx, **{**ptcore.get_wave_config(ctx), **knobs})
    frame = np.empty((H, W, 4), np.uint8)
    per_schedule = {}
    for name, schedule in (("wavefront", ptcore.SCHEDULE_WAVEFRONT), ("mega", ptcore.SCHEDULE_MEGAKERNEL)):
        tracer = ptcore.Tracer(ctx, scene, W, H, schedule)
        tracer.set_max_bounces(BOUNCES)

        def one_frame():
            tracer.trace(1)
            tracer.display_rgba8(frame)

        for _ in range(warm):
            one_frame()
        tracer.set_sample_count(0)
        ctx.synchronize()
        lat = []
        t_start = time.perf_counter()
        for _ in range(frames):
            s0 = time.perf_counter()
            one_frame()  # returns after the frame is in host memory
            lat.append((time.perf_counter() - s0) * 1e3)
        total_s = time.perf_counter() - t_start
        # the same frames without the display copy: trace, then wait for the device
        tracer.set_sample_count(0)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            tracer.trace(1)
            ctx.synchronize()
        trace_only_ms = (time.perf_counter() - t0) * 1e3 / frames
        per_schedule[name] = {"frames_per_s": frames / total_s, "ms_per_frame": total_s * 1e3 / frames,
                              "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                              "path_samples_per_s": W * H * frames / total_s,
                              "trace_only_ms_per_frame": trace_only_ms}
        print(f"[bench] interactive {name}: {per_schedule[name]}", file=sys.stderr, flush=True)
        tracer.close()
    best = max(per_schedule, key=lambda k: per_schedule[k]["frames_per_s"])
    out = {
        "metric": "interactive frames/s @1080p depth 6 (1 spp per frame + display)",
        "value": per_schedule[best]["frames_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": frames, "warmup": warm,
        "ms_per_step": per_schedule[best]["ms_per_frame"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "interactive: field-262k, 1920x1080, one sample per frame, depth 6, tone-mapped "
                               "RGBA8 frame to host each frame", "schedule": best,
                   "wave_config": ptcore.get_wave_config(ctx)},
        "schedules": per_schedule,
    }
    scene.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 5; interactive: 200 frames)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 2; interactive: 20 frames)")
    ap.add_argument("--workload", default="field",
                    choices=sorted(WORKLOADS) + ["batch", "serve", "copy", "interactive"])
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--rates", type=float, nargs="+", default=[5, 10, 20, 50, 100, 200])
    ap.add_argument("--jobs-per-rate", type=int, default=40)
    ap.add_argument("--schedule", default="wavefront", choices=["wavefront", "mega"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=5, help="CPU baseline runs of the whole job (median)")
    ap.add_argument("--traversal", type=int, default=None, choices=[0, 1, 2],
                    help="wavefront traversal: 0 batch per wave, 1 lane refill, 2 threshold refill "
                         "(default: library default)")
    ap.add_argument("--sort", type=int, default=None, choices=[0, 1, 2], help="queue sort key")
    ap.add_argument("--lds", type=int, default=None, choices=[16, 24, 32], help="LDS stack depth")
    ap.add_argument("--chunk", type=int, default=None, help="paths per wavefront chunk (default: library default)")
    ap.add_argument("--streams", type=int, default=None, choices=[0, 1, 2],
                    help="wavefront chunk lanes on separate streams (default: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pixel/tile sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import ptcore

    if args.workload == "interactive":
        print(json.dumps(bench_interactive(args)), flush=True)
        return
    args.steps = 5 if args.steps is None else args.steps
    args.warmup = 2 if args.warmup is None else args.warmup
    if args.workload in ("batch", "serve"):
        out = bench_batch(args) if args.workload == "batch" else bench_serve(args)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "copy":
        # calibration for the HBM traffic counters: a plain device copy of known size
        import torch

        n = 1 << 28  # 1 GiB of float32
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n * 4, "dispatches": 4}), flush=True)
        return
    kind_name, W, H, SPP, BOUNCES, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    world = ptcore.World.generate(getattr(ptcore, kind_name), {"field": 2, "cornell": 1, "large": 3}[args.workload], W, H)
    prep_s = time.perf_counter() - t0
    ctx = ptcore.Context(0)
    t0 = time.perf_counter()
    scene = ptcore.Scene(ctx, world.desc)
    ctx.synchronize()
    upload_s = time.perf_counter() - t0
    schedule = ptcore.SCHEDULE_WAVEFRONT if args.schedule == "wavefront" else ptcore.SCHEDULE_MEGAKERNEL
    knobs = wave_knobs(args)
    if knobs:
        ptcore.set_wave_config(ctx, **{**ptcore.get_wave_config(ctx), **knobs})
    tracer = ptcore.Tracer(ctx, scene, W, H, schedule)
    tracer.set_max_bounces(BOUNCES)
    print(f"[bench] {desc}; {world.desc.triangle_count} tris, host prep {prep_s:.2f}s, upload {upload_s:.2f}s",
          file=sys.stderr, flush=True)

    def step():
        tracer.set_sample_count(0)
        tracer.trace_all(SPP)

    # counting pass: traversal work per step (rendering is deterministic, so
    # every step issues exactly these rays and tests)
    tracer.set_profiling(2)
    step()
    counts = tracer.stats()
    print(f"[bench] counting pass: {counts}", file=sys.stderr, flush=True)

    tracer.set_profiling(0)
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # live HIP events around every traversal launch (the roofline kernel), no
    # extra synchronisation; other kernels carry no markers in the timed steps
    tracer.set_profiling(3)
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        ctx.synchronize()
        lat.append((time.perf_counter() - s0) * 1e3)
    total_s = time.perf_counter() - t_start
    ev = tracer.stats()
    # kernel-class breakdown (reported, not timed): a few more steps with every kernel's events
    breakdown_steps = min(args.steps, 3)
    tracer.set_profiling(1)
    for _ in range(breakdown_steps):
        step()
    ctx.synchronize()
    ev_all = tracer.stats()
    tracer.set_profiling(0)
    # the same steps on one lane: with two lanes every traversal launch shares
    # the GPU with the other lane's kernels, so its own duration reads longer
    ev_one = None
    if schedule == ptcore.SCHEDULE_WAVEFRONT:
        cfg0 = ptcore.get_wave_config(ctx)
        ptcore.set_wave_config(ctx, **{**cfg0, "streams": 1})
        tracer.set_profiling(3)
        for _ in range(breakdown_steps):
            step()
        ctx.synchronize()
        ev_one = tracer.stats()
        tracer.set_profiling(0)
        ptcore.set_wave_config(ctx, **cfg0)

    samples_per_step = W * H * SPP
    rays_per_step = counts["closest_rays"] + counts["shadow_rays"]
    value = samples_per_step * args.steps / total_s
    ms_per_step = total_s * 1e3 / args.steps
    result = {
        "metric": "path samples/s @1080p depth 6; rays/s; p50/p99 render-job latency",
        "value": value,
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": desc, "width": W, "height": H, "spp": SPP, "max_bounces": BOUNCES,
                   "triangles": int(world.desc.triangle_count), "schedule": args.schedule,
                   "wave_config": ptcore.get_wave_config(ctx)},
        "rays_per_s": rays_per_step * args.steps / total_s,
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
    }
    if schedule == ptcore.SCHEDULE_WAVEFRONT:
        # the traversal kernel walks the closest-hit rays of a bounce and the
        # shadow rays of the previous one; its bytes are both traversals'
        extend_launches = max(ev["extend_launches"], 1)
        bytes_per_step = (PAIR_BYTES * (counts["pair_tests"] + counts["shadow_pair_tests"]) +
                          TRI_BYTES * (counts["tri_tests"] + counts["shadow_tri_tests"]))
        launches_per_step = extend_launches / args.steps
        bytes_per_launch = bytes_per_step / launches_per_step
        avg_launch_s = ev["extend_ms"] / 1e3 / extend_launches
        achieved = bytes_per_launch / avg_launch_s / 1e9
        result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(args.workload),
                              "kernel": "k_trace (closest-hit rays of bounce b + shadow rays of bounce b-1)",
                              "avg_launch_ms": avg_launch_s * 1e3, "launches_per_step": launches_per_step,
                              "algorithmic_bytes_per_launch": bytes_per_launch}
        one_launches = max(ev_one["extend_launches"], 1)
        one_bytes = bytes_per_step / (one_launches / breakdown_steps)
        one_s = ev_one["extend_ms"] / 1e3 / one_launches
        result["roofline"]["one_lane"] = {
            "achieved": one_bytes / one_s / 1e9, "frac": one_bytes / one_s / 1e9 / HBM_PEAK_GBS,
            "avg_launch_ms": one_s * 1e3, "launches_per_step": one_launches / breakdown_steps,
            "note": "the same job on one lane in separate steps; the figures above come from the timed steps, "
                    "whose launches overlap the other lane's kernels when the render runs on two lanes"}
        result["kernel_ms_per_step"] = {"extend": ev["extend_ms"] / args.steps,
                                        "shade": ev_all["shade_ms"] / breakdown_steps,
                                        "shadow": ev_all["shadow_ms"] / breakdown_steps,
                                        "total": ev_all["total_ms"] / breakdown_steps,
                                        "note": "extend from the timed steps; the rest from separate steps with "
                                                "every kernel timed"}
    else:
        result["kernel_ms_per_step"] = {"megakernel": ev["total_ms"] / args.steps}
        bytes_per_step = PAIR_BYTES * counts["pair_tests"] + TRI_BYTES * counts["tri_tests"]
        result["roofline"] = {"bound": "hbm", "achieved": bytes_per_step / (ev["total_ms"] / 1e3 / args.steps) / 1e9,
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                              "kernel": "k_megakernel (all traversal bytes over the whole kernel)"}
        result["roofline"]["frac"] = result["roofline"]["achieved"] / HBM_PEAK_GBS
    result["traversal_counts_per_step"] = {k: counts[k] for k in ("closest_rays", "shadow_rays", "pair_tests",
                                                                   "tri_tests", "shadow_pair_tests", "shadow_tri_tests")}
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle on host cores)...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(world, W, H, BOUNCES, SPP, args.cpu_repeats)
    else:
        result["cpu_baseline"] = None
    tracer.close()
    scene.close()
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
