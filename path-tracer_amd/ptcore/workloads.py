"""Seeded synthetic workloads of the benchmark and the tests.

- the mixed-scene batch: 64 independent scenes, triangle counts log-uniform
  over 1k-500k, resolutions uniform over {256, 512, 768, 1024}^2, spp uniform
  over {4, 16, 64}, depth 6 (seed 4);
- the serving stream: Poisson arrivals of render requests over a pool of one
  4.2M-triangle scene and two 262k-triangle scenes, each request drawing its
  resolution and spp from the batch's distributions, depth 6.

bench.py and the GPU tests draw from these same functions, so a test runs
exactly the jobs the benchmark times.
"""
from __future__ import annotations

import math
import os
import random
from concurrent.futures import ThreadPoolExecutor

RESOLUTIONS = (256, 512, 768, 1024)
SPPS = (4, 16, 64)
SERVE_BOUNCES = 6
# the serving pool: (scene kind name, seed); camera aspect 1 (square requests)
SERVE_POOL = (("SCENE_LARGE", 3), ("SCENE_FIELD", 2), ("SCENE_FIELD", 12))


def mixed_jobs(seed=4, count=64):
    """The mixed-scene batch as a list of dicts (tris, width, height, spp, bounces, seed)."""
    rng = random.Random(seed)
    jobs = []
    for i in range(count):
        tris = int(math.exp(rng.uniform(math.log(1000), math.log(500000))))
        res = rng.choice(RESOLUTIONS)
        spp = rng.choice(SPPS)
        jobs.append(dict(tris=tris, width=res, height=res, spp=spp, bounces=6, seed=1000 + i))
    return jobs


def serve_requests(rate, count, pool_size=len(SERVE_POOL)):
    """`count` requests of the serving stream at `rate` jobs/s: a list of
    (scene index, resolution, spp, gap before submission in seconds)."""
    rng = random.Random(int(rate * 1000) + 5)
    reqs = [(rng.randrange(pool_size), rng.choice(RESOLUTIONS), rng.choice(SPPS)) for _ in range(count)]
    gaps = [rng.expovariate(rate) for _ in reqs]
    return [(si, res, spp, gap) for (si, res, spp), gap in zip(reqs, gaps)]


def build_worlds(specs, threads=None):
    """Prepare procedural worlds (kind, seed, width, height, triangles) in
    parallel host threads (the C++ build releases the GIL)."""
    from . import World

    def one(spec):
        kind, seed, w, h, tris = spec
        return World.generate(kind, seed, w, h, tris)

    with ThreadPoolExecutor(max_workers=threads or min(16, os.cpu_count() or 4)) as ex:
        return list(ex.map(one, specs))


def mixed_worlds(jobs):
    from . import SCENE_SIZED

    return build_worlds([(SCENE_SIZED, j["seed"], j["width"], j["height"], j["tris"]) for j in jobs])


def serve_worlds():
    from . import SCENE_FIELD, SCENE_LARGE

    kinds = {"SCENE_LARGE": SCENE_LARGE, "SCENE_FIELD": SCENE_FIELD}
    return build_worlds([(kinds[k], seed, 1024, 1024, 0) for k, seed in SERVE_POOL])
