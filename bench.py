#!/usr/bin/env python3
"""Benchmark of the path-tracing core on one MI355X.

Headline workload: the ~262k-triangle procedural field scene (a noise
heightfield sized to 261,920 triangles with 64 icospheres, 256 emissive
triangles), 1920x1080, 8 samples per pixel, max bounces 6. One step = one
render job = every sample of the frame, traced from sample 1 (the scene, its
BVH and light tables are resident in HBM before timing starts).

The default run prints one JSON line with the headline (path samples/s,
rays/s, the traversal kernel's roofline from live HIP events, the CPU
oracle's throughput on the host cores) and, beside it:
  large     the 4.2M-triangle textured scene at 3840x2160, 64 spp, depth 8:
            path samples/s and the traversal roofline (two lanes and one lane)
  batch     the 64-scene mixed batch: batched wavefront vs the plain
            megakernel and the wavefront, one scene after another
  serve     Poisson arrivals over the serving pool at 5, 50 and 200 jobs/s:
            p50/p99 submit-to-result latency and sustained path samples/s
  copy_gbs  an event-timed 1 GiB device-to-device copy (the measured HBM ceiling)

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--headline-only]
                    [--workload all|field|cornell|large|batch|serve|interactive|copy]
                    [--schedule wavefront|mega] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "path-tracer_amd"))

HBM_PEAK_GBS = 8000.0
PAIR_BYTES, TRI_BYTES = 64, 36  # a child pair of 32-B nodes; one 36-B triangle
METRIC = "path samples/s @1080p depth 6; rays/s; p50/p99 render-job latency"

WORKLOADS = {
    # name: (scene kind, seed, width, height, spp, max_bounces, description)
    "field": ("SCENE_FIELD", 2, 1920, 1080, 8, 6,
              "field-262k: 261,920-triangle noise heightfield + 64 icospheres, 256 emissive tris, 1920x1080, 8 spp, depth 6"),
    "cornell": ("SCENE_CORNELL", 1, 256, 256, 16, 8, "cornell-36: Cornell box, 256x256, 16 spp, depth 8"),
    "large": ("SCENE_LARGE", 3, 3840, 2160, 64, 8,
              "large-4.2M: 4,195,522-triangle heightfield + 4096 icospheres, 8 Perlin textures, 3840x2160, 64 spp, depth 8"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


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
        log(f"CPU baseline run {k + 1}: {runs[-1][0]:.2f} s")
    runs.sort(key=lambda r: r[0])
    dt, st = runs[len(runs) // 2]
    total = sum(r[0] for r in runs)
    what = (f"the benchmark job ({w}x{h}" if rows == h else
            f"a central band of the benchmark job's frame (rows {row0}-{row0 + rows - 1} of {w}x{h}")
    return {"value": w * rows * samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{what}, {samples} spp, depth {bounces}) run {len(runs)} times, "
                      f"median {dt:.2f} s of {total:.1f} s total; rays/s "
                      f"{((st['closest_rays'] + st['shadow_rays']) / dt):.4g}"}


def load_l1(name):
    """The L1 data-return busy fraction of the traversal from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", f"extend_l1_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("td_busy")


def load_traffic(name):
    """HBM bytes per traversal launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", f"extend_traffic_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def wave_knobs(args):
    """The wavefront configuration fields set on the command line."""
    return {k: v for k, v in (("traversal", args.traversal), ("sort_key", args.sort), ("lds_stack_depth", args.lds),
                              ("streams", args.streams), ("chunk_paths", args.chunk),
                              ("trace_waves_per_cu", args.trace_waves), ("shade_window", args.shade_window),
                              ("shade_blocks_per_cu", args.shade_blocks))
            if v is not None}


# ---------------------------------------------------------------- one render job, timed
def measure_job(ctx, name, steps, warmup, schedule_name="wavefront"):
    """Time `steps` render jobs of workload `name` (after `warmup`), with live
    HIP events around every traversal launch. Returns the measurement dict."""
    import ptcore

    kind_name, seed, W, H, SPP, BOUNCES, desc = WORKLOADS[name]
    t0 = time.perf_counter()
    world = ptcore.World.generate(getattr(ptcore, kind_name), seed, W, H)
    prep_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    scene = ptcore.Scene(ctx, world.desc)
    ctx.synchronize()
    upload_s = time.perf_counter() - t0
    schedule = ptcore.SCHEDULE_WAVEFRONT if schedule_name == "wavefront" else ptcore.SCHEDULE_MEGAKERNEL
    tracer = ptcore.Tracer(ctx, scene, W, H, schedule)
    tracer.set_max_bounces(BOUNCES)
    log(f"{desc}; {world.desc.triangle_count} tris, host prep {prep_s:.2f}s, upload {upload_s:.2f}s")

    def step():
        tracer.set_sample_count(0)
        tracer.trace_all(SPP)

    # counting pass: traversal work per step (rendering is deterministic, so
    # every step issues exactly these rays and tests)
    tracer.set_profiling(2)
    step()
    counts = tracer.stats()
    log(f"{name} counting pass: {counts}")
    tracer.set_profiling(0)
    for _ in range(warmup):
        step()
    ctx.synchronize()

    # live HIP events around every traversal launch (the roofline kernel), no
    # extra synchronisation; other kernels carry no markers in the timed steps
    tracer.set_profiling(3)
    lat = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        step()
        ctx.synchronize()
        lat.append((time.perf_counter() - s0) * 1e3)
    total_s = time.perf_counter() - t_start
    ev = tracer.stats()
    # kernel-class breakdown (reported, not timed): a few more steps with every kernel's events
    extra_steps = min(steps, 3)
    tracer.set_profiling(1)
    for _ in range(extra_steps):
        step()
    ctx.synchronize()
    ev_all = tracer.stats()
    tracer.set_profiling(0)
    # the same steps on one lane: nothing overlaps a traversal launch there
    ev_one = None
    if schedule == ptcore.SCHEDULE_WAVEFRONT:
        cfg0 = ptcore.get_wave_config(ctx)
        ptcore.set_wave_config(ctx, **{**cfg0, "streams": 1})
        tracer.set_profiling(3)
        for _ in range(extra_steps):
            step()
        ctx.synchronize()
        ev_one = tracer.stats()
        tracer.set_profiling(0)
        ptcore.set_wave_config(ctx, **cfg0)

    samples_per_step = W * H * SPP
    rays_per_step = counts["closest_rays"] + counts["shadow_rays"]
    out = {
        "workload": desc, "width": W, "height": H, "spp": SPP, "max_bounces": BOUNCES,
        "triangles": int(world.desc.triangle_count), "schedule": schedule_name,
        "value": samples_per_step * steps / total_s, "ms_per_step": total_s * 1e3 / steps,
        "rays_per_s": rays_per_step * steps / total_s,
        "step_p50_ms": percentile(lat, 50), "step_p99_ms": percentile(lat, 99),
        "host_prep_s": prep_s, "upload_s": upload_s,
        "traversal_counts_per_step": {k: counts[k] for k in ("closest_rays", "shadow_rays", "pair_tests",
                                                             "tri_tests", "shadow_pair_tests", "shadow_tri_tests")},
    }
    if schedule == ptcore.SCHEDULE_WAVEFRONT:
        # the traversal kernel walks the closest-hit rays of a bounce and the
        # shadow rays of the previous one; its bytes are both traversals'
        bytes_per_step = (PAIR_BYTES * (counts["pair_tests"] + counts["shadow_pair_tests"]) +
                          TRI_BYTES * (counts["tri_tests"] + counts["shadow_tri_tests"]))
        launches = max(ev["extend_launches"], 1)
        launches_per_step = launches / steps
        bytes_per_launch = bytes_per_step / launches_per_step
        # the union of the launches' event intervals: the time a traversal
        # kernel was running, at most the step time even where lanes overlap
        busy_s_per_step = ev["extend_busy_ms"] / 1e3 / steps
        busy_per_launch_s = busy_s_per_step / launches_per_step
        achieved = bytes_per_step / busy_s_per_step / 1e9
        traffic = load_traffic(name)
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_frac": (traffic / busy_per_launch_s / 1e9 / HBM_PEAK_GBS) if traffic else None,
                "kernel": "k_trace (closest-hit rays of bounce b + shadow rays of bounce b-1)",
                "avg_launch_ms": busy_per_launch_s * 1e3, "launches_per_step": launches_per_step,
                "algorithmic_bytes_per_launch": bytes_per_launch,
                "busy_ms_per_step": busy_s_per_step * 1e3,
                "method": "algorithmic bytes (64 per child-pair test + 36 per triangle test, counted on the device) "
                          "over the union of the traversal launches' HIP-event intervals in the timed steps",
                "overlapped_events": {
                    "avg_launch_ms": ev["extend_ms"] / launches,
                    "frac": bytes_per_launch / (ev["extend_ms"] / 1e3 / launches) / 1e9 / HBM_PEAK_GBS,
                    "note": "per-launch event durations; with two lanes they include the other lane's kernels"}}
        one_launches = max(ev_one["extend_launches"], 1)
        one_per_step = one_launches / extra_steps
        one_s = ev_one["extend_ms"] / 1e3 / one_launches
        roof["one_lane"] = {"achieved": bytes_per_step / one_per_step / one_s / 1e9,
                            "frac": bytes_per_step / one_per_step / one_s / 1e9 / HBM_PEAK_GBS,
                            "avg_launch_ms": one_s * 1e3, "launches_per_step": one_per_step,
                            "traffic_frac": (traffic / one_s / 1e9 / HBM_PEAK_GBS) if traffic else None,
                            "note": "the same job on one lane in separate steps: no launch overlaps another kernel"}
        # the traversal's nearest ceiling: the L1's rate of divergent 16-byte lane accesses
        info = ctx.info()
        accesses = (4 * (counts["pair_tests"] + counts["shadow_pair_tests"]) +
                    3 * (counts["tri_tests"] + counts["shadow_tri_tests"]))
        cu_cycles = one_s * one_per_step * info["clock_khz"] * 1e3 * info["cu_count"]
        rate = accesses / cu_cycles if cu_cycles else None
        ceiling = ctx.measure_gather(1)
        roof["one_lane"]["l1_gather"] = {
            "lane_accesses_per_cu_cycle": rate,
            "ceiling": ceiling, "frac": rate / ceiling if rate and ceiling else None,
            "ceiling_four_lanes_per_segment": ctx.measure_gather(4),
            "td_busy": load_l1(name),
            "cu_count": info["cu_count"], "clock_mhz": info["clock_khz"] / 1e3,
            "note": "lane accesses = 4 per child-pair test + 3 per triangle test (loads of at most 16 B), over "
                    "the one-lane traversal time at the peak clock; ceiling = the same rate measured on this GPU "
                    "for scattered 16-byte loads of L1-resident data, each lane its own 64-byte segment "
                    "(pt_context_measure_gather); td_busy is the L1 data-return unit's busy fraction from the "
                    "committed rocprofv3 counters"}
        out["roofline"] = roof
        out["kernel_ms_per_step"] = {"extend_busy": busy_s_per_step * 1e3,
                                     "extend_events": ev_all["extend_ms"] / extra_steps,
                                     "shade_events": ev_all["shade_ms"] / extra_steps,
                                     "total": ev_all["total_ms"] / extra_steps,
                                     "note": "extend_busy from the timed steps; the rest from separate steps with "
                                             "every kernel timed (two lanes: event sums may exceed the step)"}
    else:
        bytes_per_step = PAIR_BYTES * counts["pair_tests"] + TRI_BYTES * counts["tri_tests"]
        mk_s = ev["total_ms"] / 1e3 / steps
        out["kernel_ms_per_step"] = {"megakernel": mk_s * 1e3}
        out["roofline"] = {"bound": "hbm", "achieved": bytes_per_step / mk_s / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "traffic": None, "frac": bytes_per_step / mk_s / 1e9 / HBM_PEAK_GBS,
                           "kernel": "k_megakernel (all traversal bytes over the whole kernel)"}
    tracer.close()
    scene.close()
    return out, world


# ---------------------------------------------------------------- the mixed batch
def bench_batch(ctx, steps, warmup, count=64, autotune=True):
    """Mixed-scene batch: every job traced by the plain single-scene megakernel one after
    another, by the wavefront one after another, then all jobs as one autotuned batch."""
    import ptcore
    from ptcore import workloads

    jobs = workloads.mixed_jobs(4, count)
    t0 = time.perf_counter()
    worlds = workloads.mixed_worlds(jobs)
    prep_s = time.perf_counter() - t0
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    log(f"batch of {len(jobs)} scenes, {sum(w.desc.triangle_count for w in worlds)} triangles, "
        f"{samples / 1e6:.1f}M path samples, host prep {prep_s:.1f}s")

    def tracers(schedule):
        ts = []
        for j, s in zip(jobs, scenes):
            t = ptcore.Tracer(ctx, s, j["width"], j["height"], schedule)
            t.set_max_bounces(j["bounces"])
            ts.append(t)
        return ts

    def sequential(ts):
        ctx.synchronize()
        t0 = time.perf_counter()
        for t, j in zip(ts, jobs):
            t.set_sample_count(0)
            t.trace_all(j["spp"])
        ctx.synchronize()
        return time.perf_counter() - t0

    mk = tracers(ptcore.SCHEDULE_MEGAKERNEL)
    for t in mk[:2]:  # warm the code objects
        t.trace_all(1)
    mega_s = sequential(mk)
    for t in mk:
        t.close()
    wf = tracers(ptcore.SCHEDULE_WAVEFRONT)
    sequential(wf)  # warm the path state
    single_wf_s = sequential(wf)
    spp = [j["spp"] for j in jobs]
    cfg0 = ptcore.get_wave_config(ctx)

    def batched():
        times = []
        for k in range(warmup + steps):
            for t in wf:
                t.set_sample_count(0)
            ctx.synchronize()
            s0 = time.perf_counter()
            ptcore.trace_batch(wf, spp)
            ctx.synchronize()
            if k >= warmup:
                times.append(time.perf_counter() - s0)
        return sum(times) / len(times)

    default_s = batched()
    tune = None
    batch_s = default_s
    if autotune:
        tune = ptcore.autotune_batch(wf, spp)
        log(f"autotune: {tune}")
        batch_s = batched()
    tuned_cfg = ptcore.get_wave_config(ctx)
    ptcore.set_wave_config(ctx, **cfg0)
    for t in wf:
        t.close()
    for s in scenes:
        s.close()
    return {
        "workload": f"mixed batch: {len(jobs)} scenes, triangles log-uniform 1k-500k, resolutions "
                    f"{{256..1024}}^2, spp {{4,16,64}}, depth 6",
        "path_samples_per_step": samples, "steps": steps,
        "value": samples / batch_s, "ms_per_step": batch_s * 1e3,
        "default_config": {"value": samples / default_s, "ms_per_step": default_s * 1e3, "wave_config": cfg0},
        "tuned_config": tuned_cfg, "autotune": tune,
        "megakernel_sequential": {"value": samples / mega_s, "seconds": mega_s},
        "wavefront_sequential": {"value": samples / single_wf_s, "seconds": single_wf_s},
        "batched_vs_megakernel": mega_s / batch_s,
        "batched_vs_wavefront_sequential": single_wf_s / batch_s,
    }


# ---------------------------------------------------------------- serving
def bench_serve(ctx, rates, jobs_per_rate):
    """Serving: Poisson arrivals of render requests over resident scenes (one
    4.2M-triangle scene and two 262k-triangle scenes); each request draws its
    resolution from {256..1024}^2 and spp from {4, 16, 64}, depth 6. For each
    arrival rate: p50/p99 submit-to-result latency and sustained samples/s.
    The same requests of the highest rate traced as one batch give the
    ceiling the server is measured against."""
    import threading

    import ptcore
    from ptcore import workloads

    t0 = time.perf_counter()
    worlds = workloads.serve_worlds()
    log(f"serving pool prepared in {time.perf_counter() - t0:.1f}s")
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    results = []
    def count_for(lam):
        return jobs_per_rate or (500 if 500 / lam <= 10.0 else 200)

    for lam in rates:
        srv = ptcore.Server(ctx, hbm_budget_bytes=200 << 30, max_active_jobs=64)
        reqs = workloads.serve_requests(lam, count_for(lam))
        ids = []

        def producer():
            t_next = time.perf_counter()
            for si, res, spp, gap in reqs:
                t_next += gap
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append((srv.submit(scenes[si], res, res, spp, workloads.SERVE_BOUNCES), res, spp))

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
             "samples_per_s": samples / (last - first), "busy_samples_per_s": st["paths"] / max(st["busy_s"], 1e-9),
             "rounds": st["rounds"], "busy_s": st["busy_s"], "max_jobs_per_round": st["max_jobs_per_round"],
             "max_rounds_in_flight": st["max_rounds_in_flight"]}
        log(f"serve {r}")
        results.append(r)
    # ceiling: the highest rate's requests as one batch
    reqs = workloads.serve_requests(rates[-1], count_for(rates[-1]))

    def batch_rate(sub):
        """path samples/s and seconds of the requests `sub` traced as one batch (after a warm run)"""
        tracers = []
        for si, res, spp, _ in sub:
            t = ptcore.Tracer(ctx, scenes[si], res, res)
            t.set_max_bounces(workloads.SERVE_BOUNCES)
            tracers.append(t)
        spp = [r[2] for r in sub]
        ptcore.trace_batch(tracers, spp)  # warm
        for t in tracers:
            t.set_sample_count(0)
        ctx.synchronize()
        t0 = time.perf_counter()
        ptcore.trace_batch(tracers, spp)
        ctx.synchronize()
        secs = time.perf_counter() - t0
        for t in tracers:
            t.close()
        return sum(r[1] * r[1] * r[2] for r in sub) / secs, secs

    batch_value, batch_s = batch_rate(reqs)
    # the same requests split by scene: the 4.2M-triangle scene's and the 262k-triangle scenes'
    # requests each as one batch, and the mix rate their sample shares predict
    by_scene = {}
    shares = {}
    total_samples = sum(r[1] * r[1] * r[2] for r in reqs)
    for name, pick in (("large_4m", lambda si: si == 0), ("field_262k", lambda si: si != 0)):
        sub = [r for r in reqs if pick(r[0])]
        v, secs = batch_rate(sub)
        shares[name] = sum(r[1] * r[1] * r[2] for r in sub) / total_samples
        by_scene[name] = {"requests": len(sub), "sample_share": shares[name], "value": v, "seconds": secs}
    predicted = 1.0 / sum(shares[k] / by_scene[k]["value"] for k in by_scene)
    log(f"serve batch ceiling {batch_value / 1e9:.3f} G; by scene {by_scene}; share-weighted {predicted / 1e9:.3f} G")
    for s in scenes:
        s.close()
    top = results[-1]
    return {"workload": "serving: pool of one 4.2M-tri and two 262k-tri resident scenes; requests "
                        "{256..1024}^2, spp {4,16,64}, depth 6; 200 GB HBM budget, 64 slots; Poisson arrivals",
            "sweep": results,
            "same_requests_as_one_batch": {"value": batch_value, "seconds": batch_s,
                                           "rate_jobs_per_s": rates[-1], "by_scene": by_scene,
                                           "share_weighted_value": predicted},
            "sustained_vs_batch": top["samples_per_s"] / batch_value}


# ---------------------------------------------------------------- streaming copy
def copy_bandwidth():
    """An event-timed 1 GiB device-to-device copy in a child process (torch
    brings its own HIP runtime, kept out of this process): GB/s counts the
    bytes read plus the bytes written."""
    code = r'''
import json, torch
n = 1 << 28
a = torch.ones(n, dtype=torch.float32, device="cuda"); b = torch.empty_like(a)
for _ in range(3): b.copy_(a)
torch.cuda.synchronize()
s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
ms = []
for _ in range(10):
    s.record(); b.copy_(a); e.record(); e.synchronize(); ms.append(s.elapsed_time(e))
ms.sort()
print(json.dumps({"bytes": 2 * n * 4, "ms_median": ms[len(ms) // 2], "ms_min": ms[0]}))
'''
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        return {"value": r["bytes"] / (r["ms_median"] * 1e-3) / 1e9, "best": r["bytes"] / (r["ms_min"] * 1e-3) / 1e9,
                "unit": "GB/s", "bytes_per_copy": r["bytes"],
                "method": "torch copy_ of 1 GiB float32, HIP events, median of 10; read + written bytes"}
    except Exception as e:  # reported, never fatal to the benchmark
        return {"value": None, "error": str(e)[:200]}


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
        log(f"interactive {name}: {per_schedule[name]}")
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


def trace_union(trace_dir, bench_json):
    """From a rocprofv3 --kernel-trace of `bench.py --workload W` (CSV output
    in trace_dir) and that run's JSON line: the union of the traversal
    launches' busy intervals per timed step and the one-lane launch time, as
    the kernel trace measures them, beside the bench's own HIP-event figures.
    The non-counting k_trace dispatches come in time order as the warmup,
    timed and breakdown jobs on two lanes, then the one-lane jobs. No GPU."""
    import csv

    with open(bench_json) as f:
        bench = json.loads(f.read().strip().splitlines()[-1])
    steps, warm = bench["steps"], bench["warmup"]
    extra = min(steps, 3)
    roof = bench["roofline"]
    l2, l1 = round(roof["launches_per_step"]), round(roof["one_lane"]["launches_per_step"])
    with open(os.path.join(trace_dir, "run_kernel_trace.csv")) as f:
        rows = [r for r in csv.DictReader(f) if r["Kernel_Name"].split("(")[0].endswith(", false, false>")
                and "k_trace<" in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    iv = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows]
    if len(iv) != (warm + steps + extra) * l2 + extra * l1:
        raise SystemExit(f"unexpected traversal dispatch count {len(iv)}")

    def union(x):
        x = sorted(x)
        tot, lo, hi = 0, None, None
        for a, b in x:
            if hi is not None and a <= hi:
                hi = max(hi, b)
                continue
            if hi is not None:
                tot += hi - lo
            lo, hi = a, b
        return tot + (hi - lo if hi is not None else 0)

    timed = [iv[(warm + k) * l2:(warm + k + 1) * l2] for k in range(steps)]
    one = iv[(warm + steps + extra) * l2:]
    tc = bench["traversal_counts_per_step"]
    bytes_step = PAIR_BYTES * (tc["pair_tests"] + tc["shadow_pair_tests"]) + TRI_BYTES * (tc["tri_tests"] + tc["shadow_tri_tests"])
    busy = sum(union(j) for j in timed) / steps / 1e6
    one_ms = sum(b - a for a, b in one) / len(one) / 1e6
    return {"dispatches": len(iv), "launches_per_two_lane_job": l2, "launches_per_one_lane_job": l1,
            "algorithmic_bytes_per_step": bytes_step,
            "trace_busy_ms_per_step": busy, "trace_frac": bytes_step / (busy * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bench_busy_ms_per_step": roof["busy_ms_per_step"], "bench_frac": roof["frac"],
            "trace_per_launch_mean_ms": sum(b - a for j in timed for a, b in j) / (steps * l2) / 1e6,
            "trace_one_lane_ms": one_ms,
            "trace_one_lane_frac": bytes_step / l1 / (one_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bench_one_lane_ms": roof["one_lane"]["avg_launch_ms"], "bench_one_lane_frac": roof["one_lane"]["frac"]}


def headline_line(m, args, ctx):
    import ptcore

    return {
        "metric": METRIC, "value": m["value"], "unit": "path samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": m["workload"], "width": m["width"], "height": m["height"], "spp": m["spp"],
                   "max_bounces": m["max_bounces"], "triangles": m["triangles"], "schedule": m["schedule"],
                   "wave_config": ptcore.get_wave_config(ctx)},
        "rays_per_s": m["rays_per_s"],
        # back-to-back latency of one job; job latency under arrivals is serve.sweep[*].p50_ms/p99_ms
        "p50_ms": m["step_p50_ms"], "p99_ms": m["step_p99_ms"],
        "roofline": m["roofline"], "kernel_ms_per_step": m["kernel_ms_per_step"],
        "traversal_counts_per_step": m["traversal_counts_per_step"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 5; interactive: 200 frames)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 2; interactive: 20 frames)")
    ap.add_argument("--workload", default="all",
                    choices=["all"] + sorted(WORKLOADS) + ["batch", "serve", "copy", "interactive"])
    ap.add_argument("--headline-only", action="store_true", help="with --workload all: only the 1080p job")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--rates", type=float, nargs="+", default=[5, 50, 200])
    ap.add_argument("--jobs-per-rate", type=int, default=0,
                    help="requests per arrival rate (default: 200, and 500 where the arrivals take at most 10 s)")
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
    ap.add_argument("--trace-waves", type=int, default=None, help="traversal grid: one-wave blocks per CU")
    ap.add_argument("--shade-window", type=int, default=None, choices=[1, 2, 4],
                    help="queue rounds a shade block reorders into hits and misses")
    ap.add_argument("--shade-blocks", type=int, default=None, help="shade grid: 256-thread blocks per CU")
    ap.add_argument("--trace-union", nargs=2, metavar=("TRACE_DIR", "BENCH_JSON"),
                    help="no GPU: the traversal busy time of a rocprofv3 kernel trace of this bench")
    args = ap.parse_args()
    if args.trace_union:
        print(json.dumps(trace_union(*args.trace_union)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pixel/tile sharding across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        print(json.dumps(copy_bandwidth()), flush=True)
        return

    import ptcore

    if args.workload == "interactive":
        print(json.dumps(bench_interactive(args)), flush=True)
        return
    args.steps = 5 if args.steps is None else args.steps
    args.warmup = 2 if args.warmup is None else args.warmup
    ctx = ptcore.Context(0)
    knobs = wave_knobs(args)
    if knobs:
        ptcore.set_wave_config(ctx, **{**ptcore.get_wave_config(ctx), **knobs})
    if args.workload == "batch":
        out = bench_batch(ctx, args.steps, args.warmup, args.batch_scenes, not args.no_autotune)
        print(json.dumps({"metric": METRIC, "unit": "path samples/s", "n_gpus": 1, "dtype": "f32", **out}), flush=True)
        return
    if args.workload == "serve":
        out = bench_serve(ctx, args.rates, args.jobs_per_rate)
        print(json.dumps({"metric": "p50/p99 render-job latency under Poisson arrivals; path samples/s",
                          "unit": "ms", "n_gpus": 1, **out}), flush=True)
        return
    name = "field" if args.workload == "all" else args.workload
    m, world = measure_job(ctx, name, args.steps, args.warmup, args.schedule)
    result = headline_line(m, args, ctx)
    log(f"headline: {m['value'] / 1e9:.3f} G path samples/s, traversal frac {m['roofline']['frac']:.3f}")
    if args.workload == "all" and not args.headline_only:
        lg, _ = measure_job(ctx, "large", 2, 1, args.schedule)
        result["large"] = {k: lg[k] for k in ("workload", "value", "ms_per_step", "rays_per_s", "roofline",
                                              "traversal_counts_per_step")}
        log(f"large: {lg['value'] / 1e9:.3f} G path samples/s, traversal frac {lg['roofline']['frac']:.3f}")
        result["batch"] = bench_batch(ctx, 2, 1, args.batch_scenes, not args.no_autotune)
        log(f"batch: {result['batch']['value'] / 1e9:.3f} G path samples/s")
        result["serve"] = bench_serve(ctx, args.rates, args.jobs_per_rate)
    if args.workload == "all":
        result["copy_gbs"] = copy_bandwidth()
        log(f"copy: {result['copy_gbs']}")
    result["cpu_model"] = cpu_model()
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle on host cores)...")
        _, _, W, H, SPP, BOUNCES, _ = WORKLOADS[name]
        result["cpu_baseline"] = cpu_baseline(world, W, H, BOUNCES, SPP, args.cpu_repeats)
    else:
        result["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
