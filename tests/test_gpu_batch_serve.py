"""GPU: batching of many scenes, the autotuner knobs, and the render server.

Every per-pixel result must be bit-identical whether a job is traced alone,
in a batch with other scenes, under any wave configuration (queue sorting,
LDS stack depth, chunk size), or through the server mixed with other jobs:
the path computation is deterministic and each pixel's samples are always
accumulated in sample order.
"""
import threading
import time

import numpy as np
import pytest

import binding as oracle
import ptcore

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(ctx):
    worlds = [ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 256, 256),
              ptcore.World.generate(ptcore.SCENE_SIZED, 4, 320, 240, 20000),
              ptcore.World.generate(ptcore.SCENE_SIZED, 5, 200, 200, 120000)]
    sc = [ptcore.Scene(ctx, w.desc) for w in worlds]
    yield worlds, sc
    for s in sc:
        s.close()


JOBS = [  # (scene index, width, height, spp, bounces)
    (0, 96, 80, 4, 8), (1, 128, 96, 2, 6), (2, 64, 64, 16, 6), (1, 72, 40, 3, 3), (0, 33, 17, 5, 2),
]


def render_alone(ctx, sc, jobs):
    out = []
    for si, w, h, spp, mb in jobs:
        t = ptcore.Tracer(ctx, sc[si], w, h)
        t.set_max_bounces(mb)
        t.trace_all(spp)
        out.append(t.accumulation())
        t.close()
    return out


def test_batch_equals_individual_renders(ctx, scenes):
    _, sc = scenes
    alone = render_alone(ctx, sc, JOBS)
    tracers = []
    for si, w, h, spp, mb in JOBS:
        t = ptcore.Tracer(ctx, sc[si], w, h)
        t.set_max_bounces(mb)
        tracers.append(t)
    ptcore.trace_batch(tracers, [j[3] for j in JOBS])
    for t, ref, job in zip(tracers, alone, JOBS):
        assert t.sample_count == job[3]
        assert np.array_equal(t.accumulation(), ref), job
        t.close()


@pytest.mark.parametrize("cfg", [dict(sort_key=ptcore.SORT_OCTANT), dict(sort_key=ptcore.SORT_MATERIAL),
                                 dict(lds_stack_depth=24), dict(lds_stack_depth=32), dict(chunk_paths=1 << 16),
                                 dict(traversal=1), dict(traversal=1, sort_key=ptcore.SORT_OCTANT, lds_stack_depth=24),
                                 dict(traversal=2), dict(traversal=2, sort_key=ptcore.SORT_MATERIAL, lds_stack_depth=32,
                                                         chunk_paths=1 << 16),
                                 dict(streams=2), dict(streams=2, chunk_paths=1 << 16, sort_key=ptcore.SORT_OCTANT)])
def test_wave_configs_are_bit_identical(ctx, scenes, cfg):
    _, sc = scenes
    base = render_alone(ctx, sc, JOBS[:3])
    try:
        ptcore.set_wave_config(ctx, **cfg)
        got = render_alone(ctx, sc, JOBS[:3])
        tracers = []
        for si, w, h, spp, mb in JOBS[:3]:
            t = ptcore.Tracer(ctx, sc[si], w, h)
            t.set_max_bounces(mb)
            tracers.append(t)
        ptcore.trace_batch(tracers, [j[3] for j in JOBS[:3]])
        batch = [t.accumulation() for t in tracers]
        for t in tracers:
            t.close()
    finally:
        ptcore.set_wave_config(ctx)
    for a, b, c in zip(base, got, batch):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_automatic_large_chunks_equal_small_chunks(ctx, scenes):
    """A launch of 2^28 paths (2048x2048 at 64 spp) takes the automatic chunk
    size, which on a device with 96 GiB or more free is 2^26 paths on two
    lanes; the same render at explicit 2^24-path chunks must be identical."""
    _, sc = scenes
    w = h = 2048
    spp = 64
    assert w * h * spp == 1 << 28

    def render():
        t = ptcore.Tracer(ctx, sc[0], w, h)
        t.set_max_bounces(2)
        t.trace_all(spp)
        assert t.sample_count == spp
        out = t.accumulation()
        t.close()
        return out

    auto = render()
    try:
        ptcore.set_wave_config(ctx, chunk_paths=1 << 24)
        small = render()
    finally:
        ptcore.set_wave_config(ctx)
    assert np.isfinite(auto).all() and auto[..., :3].max() > 0.0
    assert np.array_equal(auto, small)


def test_batch_of_2p28_paths_equals_lone_renders(ctx, scenes):
    """Two jobs of 2^27 paths each: alone, each launch is below the large-chunk
    threshold (2^24-path chunks); batched, the one launch of 2^28 paths over
    two scenes takes the automatic large chunks. The results must not differ."""
    _, sc = scenes
    jobs = [(0, 2048, 1024, 64, 2), (1, 1024, 2048, 64, 2)]
    assert sum(w * h * spp for _, w, h, spp, _ in jobs) == 1 << 28
    alone = render_alone(ctx, sc, jobs)
    tracers = []
    for si, w, h, spp, mb in jobs:
        t = ptcore.Tracer(ctx, sc[si], w, h)
        t.set_max_bounces(mb)
        tracers.append(t)
    ptcore.trace_batch(tracers, [j[3] for j in jobs])
    for t, ref, job in zip(tracers, alone, jobs):
        assert t.sample_count == job[3]
        assert np.array_equal(t.accumulation(), ref), job
        t.close()


def test_batch_matches_oracle(ctx, scenes):
    worlds, sc = scenes
    si, w, h, spp, mb = JOBS[1]
    t = ptcore.Tracer(ctx, sc[si], w, h)
    t.set_max_bounces(mb)
    other = ptcore.Tracer(ctx, sc[2], 64, 64)
    ptcore.trace_batch([t, other], [spp, 4])
    ref = oracle.render_f32(worlds[si].desc, w, h, mb, spp)[0]
    g, r = t.accumulation()[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    assert np.abs(g - r).sum() / np.abs(r).sum() <= 1e-4
    assert np.array_equal(t.accumulation()[..., :3], ref[..., :3])  # and bit-exact
    t.close()
    other.close()


def test_autotune_picks_a_valid_config(ctx, scenes):
    _, sc = scenes
    tracers = []
    for si, w, h, spp, mb in JOBS[:3]:
        t = ptcore.Tracer(ctx, sc[si], w, h)
        t.set_max_bounces(mb)
        tracers.append(t)
    try:
        res = ptcore.autotune_batch(tracers, [j[3] for j in JOBS[:3]])
        assert res["trials"] >= 8
        assert res["best"]["lds_stack_depth"] in (16, 24, 32)
        assert res["best"]["sort_key"] in (0, 1, 2)
        assert res["best_ms"] <= res["initial_ms"] * 1.0001
        assert all(t.sample_count == 0 for t in tracers)
        ptcore.trace_batch(tracers, [j[3] for j in JOBS[:3]])
        got = [t.accumulation() for t in tracers]
    finally:
        ptcore.set_wave_config(ctx)
        for t in tracers:
            t.close()
    for a, b in zip(got, render_alone(ctx, sc, JOBS[:3])):
        assert np.array_equal(a, b)


def test_server_result_same_alone_and_in_the_mix(ctx, scenes):
    _, sc = scenes
    si, w, h, spp, mb = JOBS[2]
    srv = ptcore.Server(ctx, hbm_budget_bytes=8 << 30, max_active_jobs=8, round_paths=1 << 15)
    try:
        a = srv.submit(sc[si], w, h, spp, mb)
        srv.wait(a)
        alone = srv.result(a)
        # the same request among others submitted concurrently from threads
        ids = {}

        def submit(key, job):
            ids[key] = srv.submit(sc[job[0]], job[1], job[2], job[3], job[4])

        threads = [threading.Thread(target=submit, args=(k, j)) for k, j in enumerate(JOBS)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        for k, jid in ids.items():
            srv.wait(jid)
        infos = {k: srv.info(jid) for k, jid in ids.items()}
        mixed = {k: srv.result(jid) for k, jid in ids.items()}
        stats = srv.stats()
    finally:
        srv.close()
    assert np.array_equal(mixed[2], alone)
    ref = render_alone(ctx, sc, JOBS)
    for k in ids:
        assert np.array_equal(mixed[k], ref[k]), JOBS[k]
        assert infos[k]["state"] == ptcore.JOB_DONE and infos[k]["done_s"] >= infos[k]["submit_s"]
    assert stats["jobs_done"] == len(JOBS) + 1
    assert max(i["rounds"] for i in infos.values()) >= 1


def test_server_rejects_over_budget_and_cancels(ctx, scenes):
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=sc[0].footprint() + 64 * 64 * 16 + (1 << 10), max_active_jobs=1)
    try:
        with pytest.raises(ptcore.PtError) as e:
            srv.submit(sc[0], 1024, 1024, 1, 1)
        assert e.value.code == ptcore.PT_ERR_MEMORY
        assert srv.stats()["jobs_rejected"] == 1
        with pytest.raises(ptcore.PtError) as e:
            srv.submit(sc[0], 0, 16, 1, 1)
        assert e.value.code == ptcore.PT_ERR_INVALID
    finally:
        srv.close()
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=1, round_paths=1 << 12)
    try:
        first = srv.submit(sc[2], 200, 200, 16, 6)
        queued = [srv.submit(sc[0], 64, 64, 2, 2) for _ in range(3)]
        srv.cancel(queued[1])
        with pytest.raises(ptcore.PtError) as e:
            srv.wait(queued[1], 1000)
        assert e.value.code == ptcore.PT_ERR_CANCELLED
        srv.wait(first)
        srv.wait(queued[0])
        srv.wait(queued[2])
        assert srv.info(queued[1])["state"] == ptcore.JOB_CANCELLED
        with pytest.raises(ptcore.PtError) as e:
            srv.cancel(queued[1])
        assert e.value.code == ptcore.PT_ERR_STATE
        # collecting a cancelled job reports it and releases the record
        with pytest.raises(ptcore.PtError) as e:
            srv.result(queued[1])
        assert e.value.code == ptcore.PT_ERR_CANCELLED
        with pytest.raises(ptcore.PtError) as e:
            srv.info(queued[1])
        assert e.value.code == ptcore.PT_ERR_INVALID
    finally:
        srv.close()


def test_server_idle_scratch_does_not_block_admission(ctx, scenes):
    """The budget fits one job with its scene but not the path state the
    worker keeps after a round: the next job must still run once the server
    is idle."""
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=sc[0].footprint() + 64 * 64 * 16 + (1 << 16), max_active_jobs=1,
                        round_paths=1 << 12)
    try:
        a = srv.submit(sc[0], 64, 64, 2, 2)
        srv.wait(a, 30_000)
        first = srv.result(a)
        b = srv.submit(sc[0], 64, 64, 2, 2)
        srv.wait(b, 30_000)
        second = srv.result(b)
    finally:
        srv.close()
    assert np.array_equal(first, second)


def test_server_cancels_a_running_job(ctx, scenes):
    """A running job stops after its round in flight and frees its buffer;
    jobs behind it then run, and their images are unaffected."""
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=1, round_paths=1 << 12)
    try:
        big = srv.submit(sc[2], 512, 512, 64, 6)  # 16.8M paths in 4096-path rounds
        small = srv.submit(sc[0], 64, 64, 2, 2)
        deadline = time.time() + 30
        while srv.info(big)["state"] != ptcore.JOB_RUNNING and time.time() < deadline:
            time.sleep(0.001)
        assert srv.info(big)["state"] == ptcore.JOB_RUNNING
        srv.cancel(big)
        with pytest.raises(ptcore.PtError) as e:
            srv.wait(big)
        assert e.value.code == ptcore.PT_ERR_CANCELLED
        srv.wait(small)
        got = srv.result(small)
        stats = srv.stats()
    finally:
        srv.close()
    assert stats["bytes_in_use"] == 0, "the cancelled job's buffer must be released"
    assert np.array_equal(got, render_alone(ctx, sc, [(0, 64, 64, 2, 2)])[0])


def test_server_wait_times_out(ctx, scenes):
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=1, round_paths=1 << 12)
    try:
        big = srv.submit(sc[2], 256, 256, 16, 6)
        waiting = srv.submit(sc[0], 32, 32, 1, 1)
        with pytest.raises(ptcore.PtError) as e:
            srv.wait(waiting, 0)
        assert e.value.code == ptcore.PT_ERR_TIMEOUT
        srv.wait(big)
        srv.wait(waiting)
    finally:
        srv.close()


def test_server_concurrent_submit_and_cancel(ctx, scenes):
    """Threads submit jobs and cancel every other one at once, queued or
    running. Every job ends done or cancelled, collecting them all returns
    the budget to zero, and the done jobs' images are those of a lone render."""
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=4, round_paths=1 << 14)
    submitted, errors = [], []
    lock = threading.Lock()

    def worker(seed):
        rng = np.random.default_rng(seed)
        try:
            for k in range(8):
                job = (int(rng.integers(0, len(sc))), 48, 32, int(rng.choice([1, 2, 4])), 3)
                jid = srv.submit(sc[job[0]], *job[1:])
                cancel = k % 2 == 1
                if cancel:
                    try:
                        srv.cancel(jid)
                    except ptcore.PtError as e:
                        assert e.code == ptcore.PT_ERR_STATE  # it finished first
                with lock:
                    submitted.append((jid, job, cancel))
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(3)]
    try:
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        assert not errors, errors
        done = []
        for jid, job, cancel in submitted:
            try:
                srv.wait(jid, 60_000)
                done.append((job, srv.result(jid)))
            except ptcore.PtError as e:
                assert cancel and e.code == ptcore.PT_ERR_CANCELLED
                with pytest.raises(ptcore.PtError):
                    srv.result(jid)  # reports the cancellation and releases the record
        stats = srv.stats()
    finally:
        srv.close()
    assert stats["bytes_in_use"] == 0
    assert len(done) >= len(submitted) // 2
    for job, img in done[:4]:
        assert np.array_equal(img, render_alone(ctx, sc, [job])[0])


def test_server_job_matches_oracle(ctx, scenes):
    """A server job's image equals the CPU oracle's render of the same request, bit for bit."""
    worlds, sc = scenes
    si, w, h, spp, mb = JOBS[0]
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=4, round_paths=1 << 14)
    try:
        ids = [srv.submit(sc[si], w, h, spp, mb), srv.submit(sc[1], 64, 48, 3, 4)]
        for jid in ids:
            srv.wait(jid)
        got = srv.result(ids[0])
    finally:
        srv.close()
    ref = oracle.render_f32(worlds[si].desc, w, h, mb, spp)[0]
    assert np.array_equal(got[..., :3], ref[..., :3])


def test_server_shares_rounds_and_overlaps_lanes(ctx, scenes):
    """Jobs submitted together share rounds (a round holds blocks of several
    jobs) and two rounds are on the device at once; every image still equals
    its lone render."""
    _, sc = scenes
    jobs = [(2, 64, 64, 16, 6), (1, 72, 40, 3, 3), (0, 33, 17, 5, 2), (1, 128, 96, 2, 6), (2, 96, 64, 4, 6),
            (0, 96, 80, 4, 8)]
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=16, round_paths=1 << 16)
    try:
        ids = [srv.submit(sc[j[0]], *j[1:]) for j in jobs]
        for jid in ids:
            srv.wait(jid)
        got = [srv.result(jid) for jid in ids]
        stats = srv.stats()
    finally:
        srv.close()
    print(f"server stats: {stats}", flush=True)
    assert stats["max_jobs_per_round"] >= 2
    assert stats["max_rounds_in_flight"] == 2
    assert stats["rounds"] >= 3 and stats["bytes_in_use"] == 0
    for img, ref, job in zip(got, render_alone(ctx, sc, jobs), jobs):
        assert np.array_equal(img, ref), job


def test_server_refuses_scene_edits_and_destruction_while_jobs_use_it(ctx):
    """While a job over a scene is queued or running, destroying or updating
    the scene is refused (PT_ERR_STATE) and the job's image is unaffected;
    once the job is done both are allowed again."""
    world = ptcore.World.generate(ptcore.SCENE_SIZED, 21, 256, 256, 60000)
    scene = ptcore.Scene(ctx, world.desc)
    t = ptcore.Tracer(ctx, scene, 256, 256)
    t.set_max_bounces(6)
    t.trace_all(16)
    alone = t.accumulation()
    t.close()
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=2, round_paths=1 << 14)
    try:
        jid = srv.submit(scene, 256, 256, 16, 6)  # 1M paths in 16K-path rounds: many rounds
        cam = ptcore.make_camera((0.0, 1.0, 3.0))
        for call in (lambda: scene.update_camera(cam), lambda: scene.close()):
            with pytest.raises(ptcore.PtError) as e:
                call()
            assert e.value.code == ptcore.PT_ERR_STATE
        assert scene.handle  # still alive
        srv.wait(jid, 120_000)
        got = srv.result(jid)
        scene.update_camera(cam)  # allowed once the job is done
    finally:
        srv.close()
    scene.close()
    assert np.array_equal(got, alone)


def test_server_keeps_submitted_scenes_alive(ctx):
    """The Python server holds a job's scene, so dropping the caller's last
    reference while the job is pending does not free the scene under it."""
    import gc

    world = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 64, 64)
    scene = ptcore.Scene(ctx, world.desc)
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=1, round_paths=1 << 12)
    try:
        jid = srv.submit(scene, 64, 64, 8, 4)
        del scene
        gc.collect()
        srv.wait(jid)
        got = srv.result(jid)
    finally:
        srv.close()
    ref = oracle.render_f32(world.desc, 64, 64, 4, 8)[0]
    assert np.array_equal(got[..., :3], ref[..., :3])


def test_server_wait_and_collect_from_two_threads(ctx, scenes):
    """One thread waits on a job while another collects it: the waiter gets
    the result or 'already collected' (PT_ERR_INVALID), never a crash."""
    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=2, round_paths=1 << 12)
    outcomes = []
    try:
        for _ in range(8):
            jid = srv.submit(sc[0], 64, 64, 2, 2)

            def waiter():
                try:
                    srv.wait(jid, 60_000)
                    outcomes.append("ok")
                except ptcore.PtError as e:
                    outcomes.append(e.code)

            th = threading.Thread(target=waiter)
            th.start()
            while True:
                try:
                    srv.result(jid)
                    break
                except ptcore.PtError as e:
                    assert e.code == ptcore.PT_ERR_STATE
                    time.sleep(0.0005)
            th.join()
        assert srv.stats()["bytes_in_use"] == 0
    finally:
        srv.close()
    assert all(o in ("ok", ptcore.PT_ERR_INVALID) for o in outcomes), outcomes


def test_server_result_checks_the_buffer_size(ctx, scenes):
    import ctypes as C

    _, sc = scenes
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=1)
    try:
        jid = srv.submit(sc[0], 48, 32, 2, 2)
        srv.wait(jid)
        assert srv.info(jid)["result_bytes"] == 48 * 32 * 16
        small = np.empty((32, 47, 4), np.float32)
        rc = ptcore.lib().pt_server_result(srv.handle, jid, small.ctypes.data_as(C.c_void_p), small.nbytes)
        assert rc == ptcore.PT_ERR_INVALID
        assert srv.info(jid)["state"] == ptcore.JOB_DONE  # kept
        got = srv.result(jid)
    finally:
        srv.close()
    assert np.array_equal(got, render_alone(ctx, sc, [(0, 48, 32, 2, 2)])[0])


def test_server_arena_limits_admission(ctx, scenes):
    """Job buffers come from the arena reserved at creation: a request larger
    than the arena is rejected, and jobs that do not all fit at once wait for
    buffers to return and still complete exactly."""
    _, sc = scenes
    job = (1, 64, 64, 4, 3)
    arena = 2 * 64 * 64 * 16  # two jobs' buffers
    srv = ptcore.Server(ctx, hbm_budget_bytes=1 << 34, max_active_jobs=8, round_paths=1 << 13, job_arena_bytes=arena)
    try:
        assert srv.stats()["arena_bytes"] == arena
        with pytest.raises(ptcore.PtError) as e:
            srv.submit(sc[0], 128, 128, 1, 1)
        assert e.value.code == ptcore.PT_ERR_MEMORY
        ids = [srv.submit(sc[job[0]], *job[1:]) for _ in range(6)]
        got = []
        for jid in ids:
            srv.wait(jid, 60_000)
            got.append(srv.result(jid))
        assert srv.stats()["bytes_in_use"] == 0
    finally:
        srv.close()
    ref = render_alone(ctx, sc, [job])[0]
    for img in got:
        assert np.array_equal(img, ref)


ALPHA_MASK, ALPHA_BLEND = 1, 2


@pytest.fixture(scope="module")
def alpha_world(ctx):
    """The Cornell box with alpha-tested materials (a blend tall box, a mask back wall cut out and a
    mask short box that passes its cutoff). A batch or server round that holds this scene shades
    with the kernel instance that tests shadow rays inline (their any-hit walks draw random
    numbers), for every scene in the round."""
    w = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 256, 256)
    m = w.arrays()["materials"]
    for k, mode, alpha in ((5, ALPHA_BLEND, 0.55), (2, ALPHA_MASK, 0.3), (6, ALPHA_MASK, 0.8)):
        m["alpha_mode"][k] = mode
        m["base_color"][k, 3] = alpha
    s = ptcore.Scene(ctx, w.desc)
    yield w, s
    s.close()


@pytest.mark.parametrize("window", [1, 2, 4])
def test_batch_with_an_alpha_scene_equals_lone_renders(ctx, scenes, alpha_world, window):
    """Jobs on opaque scenes batched with a job on the alpha-tested scene, under each shade window:
    every job equals its lone render bit for bit, and the alpha job's tile equals the oracle."""
    aw, asc = alpha_world
    _, sc = scenes
    allsc = sc + [asc]
    jobs = JOBS[:3] + [(len(sc), 72, 56, 4, 6)]
    try:
        ptcore.set_wave_config(ctx, shade_window=window)
        alone = render_alone(ctx, allsc, jobs)
        tracers = []
        for si, w, h, spp, mb in jobs:
            t = ptcore.Tracer(ctx, allsc[si], w, h)
            t.set_max_bounces(mb)
            tracers.append(t)
        ptcore.trace_batch(tracers, [j[3] for j in jobs])
        batch = [t.accumulation() for t in tracers]
        for t in tracers:
            t.close()
    finally:
        ptcore.set_wave_config(ctx)
    for a, b, job in zip(alone, batch, jobs):
        assert np.array_equal(a, b), (window, job)
    si, w, h, spp, mb = jobs[-1]
    ref = oracle.render_f32(aw.desc, w, h, mb, spp)[0]
    assert np.array_equal(batch[-1][..., :3], ref[..., :3])


def test_server_with_an_alpha_scene(ctx, scenes, alpha_world):
    """Server rounds that mix the alpha-tested scene with opaque ones give every job its lone render."""
    _, asc = alpha_world
    _, sc = scenes
    allsc = sc + [asc]
    jobs = [(3, 64, 48, 3, 6), (1, 96, 64, 2, 5), (3, 40, 40, 5, 8), (2, 64, 64, 4, 6)]
    srv = ptcore.Server(ctx, hbm_budget_bytes=8 << 30, max_active_jobs=8, round_paths=1 << 14)
    try:
        ids = [srv.submit(allsc[si], w, h, spp, mb) for si, w, h, spp, mb in jobs]
        for jid in ids:
            srv.wait(jid)
        got = [srv.result(jid) for jid in ids]
    finally:
        srv.close()
    for g, r, job in zip(got, render_alone(ctx, allsc, jobs), jobs):
        assert np.array_equal(g, r), job
