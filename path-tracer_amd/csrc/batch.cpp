// batch.cpp - batching of many scenes, the autotuner, and the render server.
//
// No counterpart in the reference, which traces one tile of one scene per
// submission (headless.rs:207-285). Built on the job-table wavefront: a batch
// is one wavefront render whose queues hold paths of every job; the server
// runs a worker thread that admits jobs by HBM footprint and traces rounds of
// pixel blocks drawn from all active jobs (continuous batching).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pt_runtime.h"
#include "tracer.h"

using namespace ptcore;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

pt_status check_config(const pt_wave_config* c) {
    if (c->lds_stack_depth != 16 && c->lds_stack_depth != 24 && c->lds_stack_depth != 32)
        return set_error(PT_ERR_INVALID, "lds_stack_depth must be 16, 24 or 32");
    if (c->chunk_paths < 4096 || c->chunk_paths > (1u << 27))
        return set_error(PT_ERR_INVALID, "chunk_paths must be in [4096, 2^27]");
    if (c->sort_key > PT_SORT_MATERIAL) return set_error(PT_ERR_INVALID, "unknown sort key %u", c->sort_key);
    if (c->traversal > 2) return set_error(PT_ERR_INVALID, "unknown traversal mode %u", c->traversal);
    if (c->streams > 2) return set_error(PT_ERR_INVALID, "streams must be 0 (automatic), 1 or 2");
    return PT_OK;
}

pt_status batch_jobs(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples, std::vector<HostJob>& jobs) {
    if (!tracers || !samples || count == 0) return set_error(PT_ERR_INVALID, "empty batch");
    Context* c = tracers[0]->ctx;
    for (uint32_t i = 0; i < count; ++i) {
        pt_tracer* t = tracers[i];
        if (!t) return set_error(PT_ERR_INVALID, "null tracer %u", i);
        if (t->ctx != c) return set_error(PT_ERR_INVALID, "tracers of a batch must share a context");
        if (pt_status ts = tracer_require_tile(t)) return ts;
        for (uint32_t k = 0; k < i; ++k)
            if (tracers[k] == t) return set_error(PT_ERR_INVALID, "tracer %u appears twice in the batch", i);
        tracer_sync_version(t);
        if (samples[i]) jobs.push_back(tracer_job(t, samples[i]));
    }
    return PT_OK;
}

pt_status run_batch(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples) {
    std::vector<HostJob> jobs;
    pt_status st = batch_jobs(tracers, count, samples, jobs);
    if (st) return st;
    Context* c = tracers[0]->ctx;
    if (pt_status dst = use_device(c->device)) return dst;
    if (!c->batch_wave) {
        c->batch_wave = new Wavefront();
        c->batch_wave->ctx = c;
    }
    unsigned long long dummy[6] = {0};
    if ((st = c->batch_wave->render(jobs, 0, nullptr, dummy))) return st;
    for (uint32_t i = 0; i < count; ++i) tracers[i]->params.sample_count += samples[i];
    return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_context_set_wave_config(pt_context* c, const pt_wave_config* cfg) {
    if (!c || !cfg) return set_error(PT_ERR_INVALID, "null argument");
    pt_status st = check_config(cfg);
    if (st) return st;
    c->wave_cfg.lds_depth = cfg->lds_stack_depth;
    c->wave_cfg.chunk_paths = cfg->chunk_paths;
    c->wave_cfg.sort_key = cfg->sort_key;
    c->wave_cfg.traversal = cfg->traversal;
    c->wave_cfg.streams = cfg->streams;
    return PT_OK;
}

pt_status pt_context_get_wave_config(const pt_context* c, pt_wave_config* cfg) {
    if (!c || !cfg) return set_error(PT_ERR_INVALID, "null argument");
    cfg->lds_stack_depth = c->wave_cfg.lds_depth;
    cfg->chunk_paths = c->wave_cfg.chunk_paths;
    cfg->sort_key = c->wave_cfg.sort_key;
    cfg->traversal = c->wave_cfg.traversal;
    cfg->streams = c->wave_cfg.streams;
    return PT_OK;
}

pt_status pt_trace_batch(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples) {
    return run_batch(tracers, count, samples);
}

pt_status pt_autotune_batch(pt_tracer* const* tracers, uint32_t count, const uint32_t* samples, pt_tune_result* out) {
    if (!tracers || !samples || !out || count == 0) return set_error(PT_ERR_INVALID, "null argument");
    Context* c = tracers[0]->ctx;
    if (pt_status dst = use_device(c->device)) return dst;
    auto trial = [&](const ptd::WaveConfig& cfg, float& ms) -> pt_status {
        c->wave_cfg = cfg;
        for (uint32_t i = 0; i < count; ++i) tracers[i]->params.sample_count = 0;
        hipError_t se = hipStreamSynchronize(c->stream);
        if (se != hipSuccess) return hip_error(se, "autotune trial");
        auto t0 = std::chrono::steady_clock::now();
        pt_status st = run_batch(tracers, count, samples);
        if (st) return st;
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_error(e, "autotune trial");
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return PT_OK;
    };
    ptd::WaveConfig best = c->wave_cfg;
    float best_ms = 0.0f, ms = 0.0f;
    uint32_t trials = 0;
    pt_status st = trial(best, ms);  // warms allocations, then the timed initial run
    if (st == PT_OK) st = trial(best, best_ms);
    if (st) return st;
    trials = 1;
    out->initial_ms = best_ms;
    // coordinate descent: one knob at a time, keep what is faster
    const uint32_t lds_opts[] = {16, 24, 32};
    const uint32_t chunk_opts[] = {1u << 21, 1u << 22, 1u << 23, 1u << 24};
    const uint32_t sort_opts[] = {ptd::SORT_NONE, ptd::SORT_OCTANT, ptd::SORT_MATERIAL};
    const uint32_t traversal_opts[] = {ptd::TRAVERSAL_BATCH, ptd::TRAVERSAL_REFILL, ptd::TRAVERSAL_FETCH};
    const uint32_t stream_opts[] = {0, 1, 2};
    for (int knob = 0; knob < 5; ++knob) {
        const uint32_t* opts = knob == 0   ? traversal_opts
                               : knob == 1 ? chunk_opts
                               : knob == 2 ? sort_opts
                               : knob == 3 ? lds_opts
                                           : stream_opts;
        const int nopts = knob == 1 ? 4 : 3;
        ptd::WaveConfig base = best;
        for (int k = 0; k < nopts; ++k) {
            ptd::WaveConfig cand = base;
            if (knob == 0) cand.traversal = opts[k];
            else if (knob == 1) cand.chunk_paths = opts[k];
            else if (knob == 2) cand.sort_key = opts[k];
            else if (knob == 3) cand.lds_depth = opts[k];
            else cand.streams = opts[k];
            if (std::memcmp(&cand, &base, sizeof(cand)) == 0) continue;
            if ((st = trial(cand, ms))) return st;
            ++trials;
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
    }
    c->wave_cfg = best;
    for (uint32_t i = 0; i < count; ++i) tracers[i]->params.sample_count = 0;
    out->best.lds_stack_depth = best.lds_depth;
    out->best.chunk_paths = best.chunk_paths;
    out->best.sort_key = best.sort_key;
    out->best.traversal = best.traversal;
    out->best.streams = best.streams;
    out->best_ms = best_ms;
    out->trials = trials;
    return PT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ server
namespace {

struct ServerJob {
    uint64_t id;
    pt_render_request req;
    int32_t state = PT_JOB_QUEUED;
    uint32_t rounds = 0;
    double submit_s = 0, admit_s = 0, done_s = 0;
    uint64_t footprint = 0;
    float4* accum = nullptr;
    uint32_t blocks = 0;        // 8x8 blocks of the image
    uint32_t next_block = 0;    // first block not yet scheduled
    uint32_t in_flight = 0;     // blocks scheduled in the running round
    bool cancel = false;        // a running job asked to stop after its round
    std::string error;
};

}  // namespace

struct pt_server {
    Context* ctx = nullptr;
    pt_server_config cfg{};
    hipStream_t stream = nullptr;
    Wavefront wave;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, ServerJob> jobs;
    std::deque<uint64_t> queued;
    std::vector<uint64_t> active;
    uint64_t next_id = 1;
    uint64_t bytes_in_use = 0;  // admitted jobs' buffers
    uint64_t arena_bytes = 0;   // path state reserved by the worker
    bool stop = false;
    std::thread worker;
    pt_server_stats stats{};

    uint64_t job_bytes(const pt_render_request& r) const { return (uint64_t)r.width * r.height * sizeof(float4); }
    uint64_t scene_bytes_with(const Scene* extra) const {
        // resident scenes referenced by admitted jobs, each counted once
        std::vector<const Scene*> seen;
        uint64_t b = 0;
        auto add = [&](const Scene* s) {
            if (std::find(seen.begin(), seen.end(), s) != seen.end()) return;
            seen.push_back(s);
            b += s->footprint;
        };
        for (uint64_t id : active) add(jobs.at(id).req.scene);
        if (extra) add(extra);
        return b;
    }

    // whether the first queued job fits a slot and the budget now; under mu.
    // The worker's path-state scratch counts only beside running jobs: a job
    // that fits the budget alone (checked at submission) always gets to run,
    // so the queue cannot stall behind scratch that no job is using.
    bool front_fits() const {
        if (queued.empty() || active.size() >= cfg.max_active_jobs) return false;
        const ServerJob& j = jobs.at(queued.front());
        const uint64_t scratch = active.empty() ? 0 : arena_bytes;
        return bytes_in_use + job_bytes(j.req) + scene_bytes_with(j.req.scene) + scratch <= cfg.hbm_budget_bytes;
    }

    // a job leaves its slot: its buffer is released now, its record stays
    // until the caller collects it; under mu
    void finish(ServerJob& j, int32_t state, double t) {
        if (j.accum) (void)hipFree(j.accum);
        j.accum = nullptr;
        bytes_in_use -= j.footprint;
        j.footprint = 0;
        j.state = state;
        j.done_s = t;
    }

    // move queued jobs into slots while the budget allows (FIFO); under mu
    void admit() {
        while (front_fits()) {
            ServerJob& j = jobs.at(queued.front());
            const uint64_t need = job_bytes(j.req);
            hipError_t e = hipMalloc(&j.accum, need);
            if (e != hipSuccess) {
                j.accum = nullptr;
                if (!active.empty()) break;  // retry when a running job frees memory
                // nothing here will free memory: the job cannot start
                j.error = std::string("device allocation of the job buffer failed: ") + hipGetErrorString(e);
                finish(j, PT_JOB_FAILED, now_s());
                queued.pop_front();
                cv_done.notify_all();
                continue;
            }
            j.footprint = need;
            bytes_in_use += need;
            j.state = PT_JOB_RUNNING;
            j.admit_s = now_s();
            j.blocks = ((j.req.width + 7u) / 8u) * ((j.req.height + 7u) / 8u);
            active.push_back(j.id);
            queued.pop_front();
        }
    }

    // drop finished and cancelled jobs from the slots; under mu
    void retire() {
        const double t = now_s();
        bool any = false;
        for (uint64_t id : active) {
            ServerJob& j = jobs.at(id);
            if (j.state == PT_JOB_RUNNING && j.cancel && j.in_flight == 0) {
                finish(j, PT_JOB_CANCELLED, t);
                any = true;
            }
        }
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [&](uint64_t id) { return jobs.at(id).state != PT_JOB_RUNNING; }),
                     active.end());
        if (any) cv_done.notify_all();
    }

    void run() {
        (void)hipSetDevice(ctx->device);  // a failure surfaces as the first round's error
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            retire();
            admit();
            cv_work.wait(lk, [&] {
                if (stop) return true;
                for (uint64_t id : active) {
                    const ServerJob& j = jobs.at(id);
                    if (j.cancel || j.next_block < j.blocks) return true;
                }
                return front_fits();
            });
            if (stop) return;
            retire();
            admit();
            // one round: blocks from the active jobs in admission order
            const uint32_t round_paths = cfg.round_paths ? cfg.round_paths : ctx->wave_cfg.chunk_paths;
            std::vector<HostJob> round;
            std::vector<uint64_t> round_ids;
            uint64_t paths = 0;
            for (uint64_t id : active) {
                ServerJob& j = jobs.at(id);
                if (j.cancel || j.next_block >= j.blocks) continue;
                const uint64_t per_block = 64ull * j.req.samples;
                uint64_t room = paths >= round_paths ? 0 : (round_paths - paths) / per_block;
                if (room == 0 && !round.empty()) break;
                room = std::max<uint64_t>(room, 1);
                const uint32_t take = (uint32_t)std::min<uint64_t>(room, j.blocks - j.next_block);
                HostJob hj;
                hj.scene = j.req.scene;
                hj.job.scene = 0;
                hj.job.image_w = hj.job.tile_w = j.req.width;
                hj.job.image_h = hj.job.tile_h = j.req.height;
                hj.job.origin_x = hj.job.origin_y = 0;
                hj.job.max_bounces = j.req.max_bounces;
                hj.job.first_sample = 1;
                hj.job.samples = j.req.samples;
                hj.job.accum = j.accum;
                hj.job.storage = nullptr;
                hj.block_begin = j.next_block;
                hj.block_end = j.next_block + take;
                j.next_block += take;
                j.in_flight = take;
                j.rounds += 1;
                paths += per_block * take;
                round.push_back(hj);
                round_ids.push_back(id);
            }
            if (round.empty()) continue;
            lk.unlock();
            const double t0 = now_s();
            unsigned long long dummy[6] = {0};
            pt_status st = wave.render(round, 0, nullptr, dummy);
            hipError_t e = st == PT_OK ? hipStreamSynchronize(stream) : hipSuccess;
            const double t1 = now_s();
            lk.lock();
            stats.rounds += 1;
            stats.paths += paths;
            stats.busy_s += t1 - t0;
            arena_bytes = wave.arena_bytes + wave.ovf_bytes + wave.tables_bytes;
            for (uint64_t id : round_ids) {
                ServerJob& j = jobs.at(id);
                j.in_flight = 0;
                if (st != PT_OK || e != hipSuccess) {
                    j.error = st != PT_OK ? "render failed" : hipGetErrorString(e);
                    finish(j, PT_JOB_FAILED, t1);
                } else if (j.next_block >= j.blocks) {
                    j.state = PT_JOB_DONE;
                    j.done_s = t1;
                    stats.jobs_done += 1;
                }
            }
            retire();
            cv_done.notify_all();
        }
    }
};

extern "C" {

pt_status pt_server_create(pt_context* c, const pt_server_config* cfg, pt_server** out) {
    if (!c || !cfg || !out) return set_error(PT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->max_active_jobs == 0) return set_error(PT_ERR_INVALID, "max_active_jobs must be positive");
    if (cfg->hbm_budget_bytes == 0) return set_error(PT_ERR_INVALID, "hbm_budget_bytes must be positive");
    if (pt_status dst = use_device(c->device)) return dst;
    pt_server* s = new pt_server();
    s->ctx = c;
    s->cfg = *cfg;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return hip_error(e, "server stream");
    }
    s->wave.ctx = c;
    s->wave.stream_override = s->stream;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return PT_OK;
}

pt_status pt_server_destroy(pt_server* s) {
    if (!s) return PT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (uint64_t id : s->queued) s->jobs.at(id).state = PT_JOB_CANCELLED;
        s->queued.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    const hipError_t pending = hipStreamSynchronize(s->stream);
    for (auto& kv : s->jobs)
        if (kv.second.accum) (void)hipFree(kv.second.accum);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return pending == hipSuccess ? PT_OK : hip_error(pending, "work pending at server destruction");
}

pt_status pt_server_submit(pt_server* s, const pt_render_request* r, uint64_t* id) {
    if (!s || !r || !id) return set_error(PT_ERR_INVALID, "null argument");
    if (!r->scene || r->scene->ctx != s->ctx) return set_error(PT_ERR_INVALID, "request scene is not on the server's context");
    if (r->width == 0 || r->height == 0 || r->samples == 0) return set_error(PT_ERR_INVALID, "empty render request");
    if (r->max_bounces > PT_MAX_BOUNCES_LIMIT) return set_error(PT_ERR_INVALID, "max bounces above %u", PT_MAX_BOUNCES_LIMIT);
    if ((uint64_t)r->samples * 64u > (1u << 24)) return set_error(PT_ERR_INVALID, "too many samples per request");
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t need = s->job_bytes(*r) + r->scene->footprint;
    if (need > s->cfg.hbm_budget_bytes) {
        s->stats.jobs_rejected += 1;
        return set_error(PT_ERR_MEMORY, "request needs %llu bytes, over the %llu-byte budget",
                         (unsigned long long)need, (unsigned long long)s->cfg.hbm_budget_bytes);
    }
    ServerJob j;
    j.id = s->next_id++;
    j.req = *r;
    j.submit_s = now_s();
    *id = j.id;
    s->jobs.emplace(j.id, j);
    s->queued.push_back(j.id);
    s->cv_work.notify_all();
    return PT_OK;
}

pt_status pt_server_wait(pt_server* s, uint64_t id, uint32_t timeout_ms) {
    if (!s) return set_error(PT_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(PT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    auto finished = [&] {
        int st = s->jobs.at(id).state;
        return st == PT_JOB_DONE || st == PT_JOB_FAILED || st == PT_JOB_CANCELLED;
    };
    if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished))
        return set_error(PT_ERR_TIMEOUT, "job %llu not done after %u ms", (unsigned long long)id, timeout_ms);
    const ServerJob& j = s->jobs.at(id);
    if (j.state == PT_JOB_FAILED) return set_error(PT_ERR_DEVICE, "job %llu failed: %s", (unsigned long long)id, j.error.c_str());
    if (j.state == PT_JOB_CANCELLED) return set_error(PT_ERR_CANCELLED, "job %llu was cancelled", (unsigned long long)id);
    return PT_OK;
}

pt_status pt_server_info(pt_server* s, uint64_t id, pt_job_info* info) {
    if (!s || !info) return set_error(PT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(PT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    const ServerJob& j = it->second;
    info->state = j.state;
    info->rounds = j.rounds;
    info->submit_s = j.submit_s;
    info->admit_s = j.admit_s;
    info->done_s = j.done_s;
    info->footprint_bytes = j.footprint;
    return PT_OK;
}

pt_status pt_server_result(pt_server* s, uint64_t id, float* rgba) {
    if (!s) return set_error(PT_ERR_INVALID, "null argument");
    float4* buf = nullptr;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return set_error(PT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
        ServerJob& j = it->second;
        if (j.state == PT_JOB_FAILED || j.state == PT_JOB_CANCELLED) {
            // the record is released with the outcome
            const bool failed = j.state == PT_JOB_FAILED;
            const std::string why = j.error;
            s->jobs.erase(it);
            if (failed) return set_error(PT_ERR_DEVICE, "job %llu failed: %s", (unsigned long long)id, why.c_str());
            return set_error(PT_ERR_CANCELLED, "job %llu was cancelled", (unsigned long long)id);
        }
        if (j.state != PT_JOB_DONE) return set_error(PT_ERR_STATE, "job %llu is not finished", (unsigned long long)id);
        if (!rgba) return set_error(PT_ERR_INVALID, "null argument");
        buf = j.accum;
        bytes = j.footprint;
        j.accum = nullptr;
        s->bytes_in_use -= j.footprint;
        s->jobs.erase(it);
    }
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e == hipSuccess) e = hipMemcpy(rgba, buf, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    s->cv_work.notify_all();  // memory freed: more jobs may fit
    return e == hipSuccess ? PT_OK : hip_error(e, "job result copy");
}

pt_status pt_server_cancel(pt_server* s, uint64_t id) {
    if (!s) return set_error(PT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(PT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    ServerJob& j = it->second;
    if (j.state == PT_JOB_QUEUED) {
        j.state = PT_JOB_CANCELLED;
        j.done_s = now_s();
        s->queued.erase(std::remove(s->queued.begin(), s->queued.end(), id), s->queued.end());
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // the next queued job may now be first in line
        return PT_OK;
    }
    if (j.state != PT_JOB_RUNNING) return set_error(PT_ERR_STATE, "job %llu has finished", (unsigned long long)id);
    // the worker stops scheduling it and releases it after its round in flight
    j.cancel = true;
    s->cv_work.notify_all();
    return PT_OK;
}

pt_status pt_server_stats_get(pt_server* s, pt_server_stats* out) {
    if (!s || !out) return set_error(PT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->bytes_in_use = s->bytes_in_use;
    return PT_OK;
}

}  // extern "C"
