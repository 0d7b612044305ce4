"""GPU: the mixed-scene batch and the serving stream at their benchmark sizes.

- Batch: the 64 seeded scenes of the benchmark, drawn by the same function
  (triangle counts log-uniform over 1k-500k, resolutions {256..1024}^2,
  spp {4, 16, 64}, depth 6). Every job's accumulation from one batched launch
  sequence equals its lone wavefront render bit for bit (the full-size
  property). Tiles of the smallest, median and largest scene equal the CPU
  oracle.
- Serving: one 4.2M-triangle scene and two 262k-triangle scenes resident,
  requests drawn as the benchmark draws them (Poisson arrivals, resolution
  and spp from the batch's distributions, depth 6). Every result equals its
  lone render, one tile per scene equals the oracle, jobs shared rounds and
  two rounds were on the device at once.
"""
import time

import numpy as np
import pytest

import binding as oracle
import ptcore
from ptcore import workloads
from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu


def lone_render(ctx, scene, w, h, spp, bounces):
    t = ptcore.Tracer(ctx, scene, w, h)
    t.set_max_bounces(bounces)
    t.trace_all(spp)
    img = t.accumulation()
    t.close()
    return img


def oracle_tile(world, w, h, bounces, spp, tw=32, th=32):
    origin = ((w - tw) // 2, (h - th) // 2)
    ref, _, _ = oracle.render_f32(world.desc, w, h, bounces, spp, tile=(tw, th), origin=origin)
    return ref, origin


@pytest.fixture(scope="module")
def batch(ctx):
    jobs = workloads.mixed_jobs(4, 64)
    t0 = time.perf_counter()
    worlds = workloads.mixed_worlds(jobs)
    print(f"mixed batch: {len(jobs)} scenes, {sum(w.desc.triangle_count for w in worlds)} triangles, "
          f"prepared in {time.perf_counter() - t0:.1f}s", flush=True)
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    yield jobs, worlds, scenes
    for s in scenes:
        s.close()


def test_mixed_batch_equals_lone_renders(ctx, batch):
    jobs, worlds, scenes = batch
    tracers = []
    for j, s in zip(jobs, scenes):
        t = ptcore.Tracer(ctx, s, j["width"], j["height"])
        t.set_max_bounces(j["bounces"])
        tracers.append(t)
    ptcore.trace_batch(tracers, [j["spp"] for j in jobs])
    mism = []
    for k, (t, j, s) in enumerate(zip(tracers, jobs, scenes)):
        got = t.accumulation()
        t.close()
        if not np.array_equal(got, lone_render(ctx, s, j["width"], j["height"], j["spp"], j["bounces"])):
            mism.append(k)
    print(f"mixed batch: {len(jobs)} jobs, {sum(j['width'] * j['height'] * j['spp'] for j in jobs) / 1e6:.0f}M "
          f"path samples, jobs differing from their lone render: {mism}", flush=True)
    assert not mism


@pytest.mark.parametrize("which", ["smallest", "median", "largest"])
def test_mixed_batch_tiles_match_oracle(ctx, batch, which):
    jobs, worlds, scenes = batch
    order = sorted(range(len(jobs)), key=lambda k: jobs[k]["tris"])
    k = {"smallest": order[0], "median": order[len(order) // 2], "largest": order[-1]}[which]
    j = jobs[k]
    others = [i for i in order if i != k][:3]  # batched with three other scenes
    tracers = []
    for i in [k] + others:
        t = ptcore.Tracer(ctx, scenes[i], jobs[i]["width"], jobs[i]["height"])
        t.set_max_bounces(jobs[i]["bounces"])
        tracers.append(t)
    ptcore.trace_batch(tracers, [jobs[i]["spp"] for i in [k] + others])
    got = tracers[0].accumulation()
    for t in tracers:
        t.close()
    ref, (ox, oy) = oracle_tile(worlds[k], j["width"], j["height"], j["bounces"], j["spp"], 64, 64)
    print(f"{which} scene: {j}", flush=True)
    assert_parity(got[oy:oy + 64, ox:ox + 64], ref, f"batch {which} scene tile")


@pytest.fixture(scope="module")
def serve_pool(ctx):
    t0 = time.perf_counter()
    worlds = workloads.serve_worlds()
    print(f"serving pool prepared in {time.perf_counter() - t0:.1f}s", flush=True)
    scenes = [ptcore.Scene(ctx, w.desc) for w in worlds]
    yield worlds, scenes
    for s in scenes:
        s.close()


def test_serving_stream_results(ctx, serve_pool):
    worlds, scenes = serve_pool
    rate, count = 200.0, 24
    reqs = workloads.serve_requests(rate, count)
    srv = ptcore.Server(ctx, hbm_budget_bytes=200 << 30, max_active_jobs=64)
    ids = []
    try:
        # the first requests arrive together, the rest with Poisson gaps
        t_next = time.perf_counter()
        for n, (si, res, spp, gap) in enumerate(reqs):
            if n >= 4:
                t_next += gap
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            ids.append(srv.submit(scenes[si], res, res, spp, workloads.SERVE_BOUNCES))
        results = []
        for jid in ids:
            srv.wait(jid, 600_000)
            info = srv.info(jid)
            results.append((srv.result(jid), info))
        stats = srv.stats()
    finally:
        srv.close()
    print(f"serving: {count} jobs, stats {stats}", flush=True)
    assert stats["max_jobs_per_round"] >= 2, "no round held blocks of two jobs"
    assert stats["max_rounds_in_flight"] == 2, "rounds never overlapped on the device"
    assert stats["jobs_done"] == count and stats["bytes_in_use"] == 0
    mism = []
    for k, ((si, res, spp, _), (img, info)) in enumerate(zip(reqs, results)):
        assert info["state"] == ptcore.JOB_DONE and info["done_s"] >= info["admit_s"] >= info["submit_s"]
        if not np.array_equal(img, lone_render(ctx, scenes[si], res, res, spp, workloads.SERVE_BOUNCES)):
            mism.append(k)
    assert not mism, f"server results differing from lone renders: {mism}"
    # one tile per scene against the oracle
    for si in range(len(scenes)):
        k = next(k for k, r in enumerate(reqs) if r[0] == si)
        _, res, spp, _ = reqs[k]
        ref, (ox, oy) = oracle_tile(worlds[si], res, res, workloads.SERVE_BOUNCES, spp)
        assert_parity(results[k][0][oy:oy + 32, ox:ox + 32], ref, f"serving scene {si} request {k}")
