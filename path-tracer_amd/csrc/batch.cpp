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
    if (c->chunk_paths != 0 && (c->chunk_paths < 4096 || c->chunk_paths > (1u << 27)))
        return set_error(PT_ERR_INVALID, "chunk_paths must be 0 (automatic) or in [4096, 2^27]");
    if (c->sort_key > PT_SORT_MATERIAL) return set_error(PT_ERR_INVALID, "unknown sort key %u", c->sort_key);
    if (c->traversal > 2) return set_error(PT_ERR_INVALID, "unknown traversal mode %u", c->traversal);
    if (c->streams > 2) return set_error(PT_ERR_INVALID, "streams must be 0 (automatic), 1 or 2");
    if (c->trace_waves_per_cu < 16 || c->trace_waves_per_cu > 2048)
        return set_error(PT_ERR_INVALID, "trace_waves_per_cu must be in [16, 2048]");
    if (c->shade_blocks_per_cu < 1 || c->shade_blocks_per_cu > 256)
        return set_error(PT_ERR_INVALID, "shade_blocks_per_cu must be in [1, 256]");
    if (c->shade_window != 0 && c->shade_window != 1 && c->shade_window != 2 && c->shade_window != 4)
        return set_error(PT_ERR_INVALID, "shade_window must be 0 (automatic), 1, 2 or 4");
    if (c->lane_split_paths == 0) return set_error(PT_ERR_INVALID, "lane_split_paths must be positive");
    return PT_OK;
}

ptd::WaveConfig from_api(const pt_wave_config& c) {
    ptd::WaveConfig w;
    w.lds_depth = c.lds_stack_depth;
    w.chunk_paths = c.chunk_paths;
    w.sort_key = c.sort_key;
    w.traversal = c.traversal;
    w.streams = c.streams;
    w.trace_waves_per_cu = c.trace_waves_per_cu;
    w.shade_blocks_per_cu = c.shade_blocks_per_cu;
    w.shade_window = c.shade_window;
    w.lane_split_paths = c.lane_split_paths;
    return w;
}

pt_wave_config to_api(const ptd::WaveConfig& w) {
    pt_wave_config c;
    c.lds_stack_depth = w.lds_depth;
    c.chunk_paths = w.chunk_paths;
    c.sort_key = w.sort_key;
    c.traversal = w.traversal;
    c.streams = w.streams;
    c.trace_waves_per_cu = w.trace_waves_per_cu;
    c.shade_blocks_per_cu = w.shade_blocks_per_cu;
    c.shade_window = w.shade_window;
    c.lane_split_paths = w.lane_split_paths;
    return c;
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
    c->wave_cfg = from_api(*cfg);
    return PT_OK;
}

pt_status pt_context_get_wave_config(const pt_context* c, pt_wave_config* cfg) {
    if (!c || !cfg) return set_error(PT_ERR_INVALID, "null argument");
    *cfg = to_api(c->wave_cfg);
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
    if (st) {
        c->wave_cfg = best;
        return st;
    }
    trials = 1;
    out->initial_ms = best_ms;
    // coordinate descent: one knob at a time, keep what is faster
    struct Knob {
        uint32_t ptd::WaveConfig::*field;
        std::vector<uint32_t> opts;
    };
    const Knob knobs[] = {
        {&ptd::WaveConfig::traversal, {ptd::TRAVERSAL_BATCH, ptd::TRAVERSAL_REFILL, ptd::TRAVERSAL_FETCH}},
        {&ptd::WaveConfig::chunk_paths, {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 26}},
        {&ptd::WaveConfig::sort_key, {ptd::SORT_NONE, ptd::SORT_OCTANT, ptd::SORT_MATERIAL}},
        {&ptd::WaveConfig::lds_depth, {16, 24, 32}},
        {&ptd::WaveConfig::streams, {0, 1, 2}},
        {&ptd::WaveConfig::trace_waves_per_cu, {256, 512, 1024}},
        {&ptd::WaveConfig::shade_blocks_per_cu, {16, 32, 64}},
        {&ptd::WaveConfig::shade_window, {1, 2, 4}},
    };
    for (const Knob& knob : knobs) {
        const ptd::WaveConfig base = best;
        for (uint32_t v : knob.opts) {
            ptd::WaveConfig cand = base;
            cand.*knob.field = v;
            if (std::memcmp(&cand, &base, sizeof(cand)) == 0) continue;
            st = trial(cand, ms);
            if (st == PT_ERR_MEMORY) continue;  // too large for the device now: not a candidate
            if (st) {
                c->wave_cfg = best;
                for (uint32_t i = 0; i < count; ++i) tracers[i]->params.sample_count = 0;
                return st;
            }
            ++trials;
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
    }
    c->wave_cfg = best;
    for (uint32_t i = 0; i < count; ++i) tracers[i]->params.sample_count = 0;
    out->best = to_api(best);
    out->best_ms = best_ms;
    out->trials = trials;
    return PT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ server
namespace {

// First-fit sub-allocator over one device range reserved at server creation.
// Blocks are 256-byte aligned; freed blocks merge with free neighbours.
struct JobArena {
    char* base = nullptr;
    uint64_t bytes = 0;
    std::map<uint64_t, uint64_t> free_blocks;  // offset -> size
    static uint64_t round_up(uint64_t n) { return (n + 255) & ~(uint64_t)255; }
    void reset(char* b, uint64_t n) {
        base = b;
        bytes = n & ~(uint64_t)255;
        free_blocks.clear();
        if (bytes) free_blocks[0] = bytes;
    }
    // nullptr when no free block is large enough
    bool can_take(uint64_t n) const {
        n = round_up(n);
        for (const auto& kv : free_blocks)
            if (kv.second >= n) return true;
        return false;
    }
    void* take(uint64_t n) {
        n = round_up(n);
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < n) continue;
            const uint64_t off = it->first, size = it->second;
            free_blocks.erase(it);
            if (size > n) free_blocks[off + n] = size - n;
            return base + off;
        }
        return nullptr;
    }
    void give(void* p, uint64_t n) {
        if (!p) return;
        n = round_up(n);
        uint64_t off = (uint64_t)((char*)p - base);
        auto next = free_blocks.lower_bound(off);
        if (next != free_blocks.end() && off + n == next->first) {
            n += next->second;
            next = free_blocks.erase(next);
        }
        if (next != free_blocks.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_blocks[off] = n;
    }
};

struct ServerJob {
    uint64_t id;
    pt_render_request req;
    int32_t state = PT_JOB_QUEUED;
    uint32_t rounds = 0;
    double submit_s = 0, admit_s = 0, done_s = 0;
    uint64_t footprint = 0;
    float4* accum = nullptr;    // in the server's job arena
    uint32_t blocks = 0;        // 8x8 blocks of the image
    uint32_t next_block = 0;    // first block not yet scheduled
    uint32_t in_flight = 0;     // rounds on the device that trace blocks of this job
    bool cancel = false;        // a running job asked to stop after its rounds in flight
    bool failed = false;        // a round of the job failed; it ends once no round is in flight
    bool holds_scene = false;   // counted in req.scene->server_jobs
    std::string error;
};

// A round on the device: launched on `lane`, complete when `done` has fired.
struct Round {
    int lane;
    std::vector<uint64_t> ids;
    uint64_t paths;
    pt_status status;  // of the launches
};

constexpr int SERVER_LANES = 2;

}  // namespace

struct pt_server {
    Context* ctx = nullptr;
    pt_server_config cfg{};
    // two lanes with their own path state on the context's two streams: rounds
    // alternate between them, and the next round is launched while the current
    // one runs
    hipStream_t lane_stream[SERVER_LANES] = {nullptr, nullptr};
    hipEvent_t lane_done[SERVER_LANES] = {nullptr, nullptr};
    bool lane_busy[SERVER_LANES] = {false, false};
    Wavefront wave[SERVER_LANES];
    void* arena_mem = nullptr;
    JobArena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, ServerJob> jobs;
    std::deque<uint64_t> queued;
    std::vector<uint64_t> active;
    std::deque<Round> in_flight;  // oldest first
    uint64_t next_id = 1;
    uint64_t bytes_in_use = 0;  // admitted jobs' buffers
    uint64_t scratch_bytes = 0; // path state reserved by the worker's lanes
    double busy_since = 0;      // start of the current stretch with rounds in flight
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
        const uint64_t scratch = active.empty() ? 0 : scratch_bytes;
        return bytes_in_use + job_bytes(j.req) + scene_bytes_with(j.req.scene) + scratch <= cfg.hbm_budget_bytes;
    }

    // the first queued job fits now and the arena has a buffer for it; under mu
    bool front_admissible() const { return front_fits() && arena.can_take(job_bytes(jobs.at(queued.front()).req)); }

    // the job no longer uses its scene; under mu
    void release_scene(ServerJob& j) {
        if (j.holds_scene) {
            j.req.scene->server_jobs.fetch_sub(1);
            j.holds_scene = false;
        }
    }

    // a job leaves its slot: its buffer returns to the arena now, its record
    // stays until the caller collects it; under mu, with no round in flight for it
    void finish(ServerJob& j, int32_t state, double t) {
        arena.give(j.accum, j.footprint);
        j.accum = nullptr;
        bytes_in_use -= j.footprint;
        j.footprint = 0;
        j.state = state;
        j.done_s = t;
        release_scene(j);
    }

    // move queued jobs into slots while the budget and the arena allow (FIFO); under mu
    void admit() {
        while (front_admissible()) {
            ServerJob& j = jobs.at(queued.front());
            const uint64_t need = job_bytes(j.req);
            void* buf = arena.take(need);
            j.accum = (float4*)buf;
            j.footprint = need;
            bytes_in_use += need;
            j.state = PT_JOB_RUNNING;
            j.admit_s = now_s();
            j.blocks = ((j.req.width + 7u) / 8u) * ((j.req.height + 7u) / 8u);
            active.push_back(j.id);
            queued.pop_front();
        }
    }

    // end jobs whose last round came back, and drop finished jobs from the slots; under mu
    void retire() {
        const double t = now_s();
        bool any = false;
        for (uint64_t id : active) {
            ServerJob& j = jobs.at(id);
            if (j.state != PT_JOB_RUNNING || j.in_flight != 0) continue;
            if (j.failed) {
                finish(j, PT_JOB_FAILED, t);
                any = true;
            } else if (j.next_block >= j.blocks) {
                // every block traced and back: the buffer stays until collected
                j.state = PT_JOB_DONE;
                j.done_s = t;
                stats.jobs_done += 1;
                release_scene(j);
                any = true;
            } else if (j.cancel) {
                finish(j, PT_JOB_CANCELLED, t);
                any = true;
            }
        }
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [&](uint64_t id) { return jobs.at(id).state != PT_JOB_RUNNING; }),
                     active.end());
        if (any) cv_done.notify_all();
    }

    // an active job that retire() would end now (cancelled or failed, nothing in flight); under mu
    bool retire_pending() const {
        for (uint64_t id : active) {
            const ServerJob& j = jobs.at(id);
            if (j.in_flight == 0 && (j.cancel || j.failed)) return true;
        }
        return false;
    }
    bool schedulable() const {
        for (uint64_t id : active) {
            const ServerJob& j = jobs.at(id);
            if (!j.cancel && !j.failed && j.next_block < j.blocks) return true;
        }
        return false;
    }
    int free_lane() const {
        for (int l = 0; l < SERVER_LANES; ++l)
            if (!lane_busy[l]) return l;
        return -1;
    }

    // blocks from the active jobs in admission order, up to the round size; under mu
    Round build_round(int lane, std::vector<HostJob>& round) {
        Round r{lane, {}, 0, PT_OK};
        const uint32_t round_paths =
            cfg.round_paths ? cfg.round_paths : ctx->wave_cfg.chunk_paths ? ctx->wave_cfg.chunk_paths : ptd::CHUNK_DEFAULT;
        for (uint64_t id : active) {
            ServerJob& j = jobs.at(id);
            if (j.cancel || j.failed || j.next_block >= j.blocks) continue;
            const uint64_t per_block = 64ull * j.req.samples;
            uint64_t room = r.paths >= round_paths ? 0 : (round_paths - r.paths) / per_block;
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
            j.in_flight += 1;
            j.rounds += 1;
            r.paths += per_block * take;
            round.push_back(hj);
            r.ids.push_back(id);
        }
        return r;
    }

    // a round came back (or failed); under mu
    void complete(const Round& r, bool ok, const std::string& why) {
        const double t = now_s();
        stats.rounds += 1;
        stats.paths += r.paths;
        lane_busy[r.lane] = false;
        for (uint64_t id : r.ids) {
            ServerJob& j = jobs.at(id);
            j.in_flight -= 1;
            if (!ok && !j.failed) {
                j.failed = true;
                j.error = why;
            }
        }
        if (in_flight.empty()) stats.busy_s += t - busy_since;
    }

    void run() {
        (void)hipSetDevice(ctx->device);  // a failure surfaces as the first round's error
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // rounds that came back, oldest first (they may end out of order)
            for (size_t k = 0; k < in_flight.size();) {
                Round& r = in_flight[k];
                hipError_t q = r.status == PT_OK ? hipEventQuery(lane_done[r.lane]) : hipSuccess;
                if (q == hipErrorNotReady) {
                    ++k;
                    continue;
                }
                Round done = std::move(r);
                in_flight.erase(in_flight.begin() + (ptrdiff_t)k);
                std::string why;
                if (done.status != PT_OK) why = "render failed";
                else if (q != hipSuccess) why = hipGetErrorString(q);
                complete(done, why.empty(), why);
            }
            retire();
            admit();
            if (stop && in_flight.empty()) return;
            const int lane = stop ? -1 : free_lane();
            if (lane >= 0 && schedulable()) {
                std::vector<HostJob> round;
                Round r = build_round(lane, round);
                stats.max_jobs_per_round = std::max<uint32_t>(stats.max_jobs_per_round, (uint32_t)round.size());
                lane_busy[lane] = true;
                if (in_flight.empty()) busy_since = now_s();
                lk.unlock();
                unsigned long long dummy[6] = {0};
                pt_status st = wave[lane].render(round, 0, nullptr, dummy);
                hipError_t e = hipSuccess;
                if (st == PT_OK) e = hipEventRecord(lane_done[lane], lane_stream[lane]);
                if (st != PT_OK || e != hipSuccess) {
                    // some launches may be queued: let them drain before the buffers are reused
                    (void)hipStreamSynchronize(lane_stream[lane]);
                    if (st == PT_OK) st = PT_ERR_DEVICE;
                }
                lk.lock();
                r.status = st;
                in_flight.push_back(std::move(r));
                stats.max_rounds_in_flight = std::max<uint32_t>(stats.max_rounds_in_flight, (uint32_t)in_flight.size());
                scratch_bytes = wave[0].device_bytes() + wave[1].device_bytes();
                continue;  // maybe launch on the other lane too
            }
            if (in_flight.empty()) {
                cv_work.wait(lk, [&] { return stop || schedulable() || front_admissible() || retire_pending(); });
                continue;
            }
            if (lane < 0) {
                // both lanes busy (or stopping): wait for the oldest round
                const int l = in_flight.front().lane;
                lk.unlock();
                (void)hipEventSynchronize(lane_done[l]);  // its result is read by the query above
                lk.lock();
                continue;
            }
            // a lane is free but nothing can use it yet: wake for new work or to check the rounds
            cv_work.wait_for(lk, std::chrono::microseconds(200),
                             [&] { return stop || schedulable() || front_admissible() || retire_pending(); });
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
    uint64_t arena = cfg->job_arena_bytes ? cfg->job_arena_bytes : std::min<uint64_t>(cfg->hbm_budget_bytes, 4ull << 30);
    arena = std::max<uint64_t>(arena & ~(uint64_t)255, 256);
    s->cfg.job_arena_bytes = arena;
    hipError_t e = hipMalloc(&s->arena_mem, arena);
    s->lane_stream[0] = c->stream;
    s->lane_stream[1] = c->lane_stream;
    for (int l = 0; l < SERVER_LANES && e == hipSuccess; ++l) {
        e = hipEventCreateWithFlags(&s->lane_done[l], hipEventDisableTiming);
        s->wave[l].ctx = c;
        s->wave[l].stream_override = s->lane_stream[l];
        s->wave[l].one_lane = true;
    }
    if (e != hipSuccess) {
        for (int l = 0; l < SERVER_LANES; ++l) {
            if (s->lane_done[l]) (void)hipEventDestroy(s->lane_done[l]);
        }
        if (s->arena_mem) (void)hipFree(s->arena_mem);
        delete s;
        return hip_error(e, "server arena and lanes");
    }
    s->arena.reset((char*)s->arena_mem, arena);
    s->stats.arena_bytes = s->arena.bytes;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return PT_OK;
}

pt_status pt_server_destroy(pt_server* s) {
    if (!s) return PT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (uint64_t id : s->queued) {
            ServerJob& j = s->jobs.at(id);
            j.state = PT_JOB_CANCELLED;
            s->release_scene(j);
        }
        s->queued.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();  // the worker returns once no round is in flight
    (void)hipSetDevice(s->ctx->device);
    hipError_t pending = hipSuccess;
    for (int l = 0; l < SERVER_LANES; ++l) {
        const hipError_t e = hipStreamSynchronize(s->lane_stream[l]);
        if (pending == hipSuccess) pending = e;
    }
    for (auto& kv : s->jobs) s->release_scene(kv.second);
    for (int l = 0; l < SERVER_LANES; ++l) {
        (void)hipEventDestroy(s->lane_done[l]);
    }
    (void)hipFree(s->arena_mem);
    delete s;  // the lanes' path state is released here, with their work finished
    return pending == hipSuccess ? PT_OK : hip_error(pending, "work pending at server destruction");
}

pt_status pt_server_submit(pt_server* s, const pt_render_request* r, uint64_t* id) {
    if (!s || !r || !id) return set_error(PT_ERR_INVALID, "null argument");
    if (!r->scene || r->scene->ctx != s->ctx) return set_error(PT_ERR_INVALID, "request scene is not on the server's context");
    if (r->width == 0 || r->height == 0 || r->samples == 0) return set_error(PT_ERR_INVALID, "empty render request");
    if (r->max_bounces > PT_MAX_BOUNCES_LIMIT) return set_error(PT_ERR_INVALID, "max bounces above %u", PT_MAX_BOUNCES_LIMIT);
    if ((uint64_t)r->samples * 64u > (1u << 24)) return set_error(PT_ERR_INVALID, "too many samples per request");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(PT_ERR_STATE, "server is shutting down");
    const uint64_t buf = s->job_bytes(*r);
    const uint64_t need = buf + r->scene->footprint;
    if (need > s->cfg.hbm_budget_bytes || JobArena::round_up(buf) > s->arena.bytes) {
        s->stats.jobs_rejected += 1;
        if (need > s->cfg.hbm_budget_bytes)
            return set_error(PT_ERR_MEMORY, "request needs %llu bytes, over the %llu-byte budget",
                             (unsigned long long)need, (unsigned long long)s->cfg.hbm_budget_bytes);
        return set_error(PT_ERR_MEMORY, "request's %llu-byte buffer exceeds the %llu-byte job arena",
                         (unsigned long long)buf, (unsigned long long)s->arena.bytes);
    }
    ServerJob j;
    j.id = s->next_id++;
    j.req = *r;
    j.submit_s = now_s();
    j.holds_scene = true;
    r->scene->server_jobs.fetch_add(1);
    *id = j.id;
    s->jobs.emplace(j.id, j);
    s->queued.push_back(j.id);
    s->cv_work.notify_all();
    return PT_OK;
}

pt_status pt_server_wait(pt_server* s, uint64_t id, uint32_t timeout_ms) {
    if (!s) return set_error(PT_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->jobs.find(id) == s->jobs.end()) return set_error(PT_ERR_INVALID, "unknown job %llu", (unsigned long long)id);
    // another thread may collect the job meanwhile: a missing record ends the wait
    auto finished = [&] {
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) return true;
        const int st = it->second.state;
        return st == PT_JOB_DONE || st == PT_JOB_FAILED || st == PT_JOB_CANCELLED;
    };
    if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished))
        return set_error(PT_ERR_TIMEOUT, "job %llu not done after %u ms", (unsigned long long)id, timeout_ms);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end())
        return set_error(PT_ERR_INVALID, "job %llu was already collected", (unsigned long long)id);
    const ServerJob& j = it->second;
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
    info->result_bytes = s->job_bytes(j.req);
    return PT_OK;
}

pt_status pt_server_result(pt_server* s, uint64_t id, float* rgba, size_t capacity) {
    if (!s) return set_error(PT_ERR_INVALID, "null argument");
    float4* buf = nullptr;
    uint64_t bytes = 0, held = 0;
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
        bytes = s->job_bytes(j.req);
        if (capacity < bytes)
            return set_error(PT_ERR_INVALID, "result of job %llu needs %llu bytes, the buffer holds %zu",
                             (unsigned long long)id, (unsigned long long)bytes, capacity);
        // the record leaves now; the buffer returns to the arena after the copy
        buf = j.accum;
        held = j.footprint;
        j.accum = nullptr;
        s->jobs.erase(it);
    }
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e == hipSuccess) e = hipMemcpy(rgba, buf, bytes, hipMemcpyDeviceToHost);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->arena.give(buf, held);
        s->bytes_in_use -= held;
    }
    s->cv_work.notify_all();  // memory returned: more jobs may fit
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
        s->release_scene(j);
        s->queued.erase(std::remove(s->queued.begin(), s->queued.end(), id), s->queued.end());
        s->cv_done.notify_all();
        s->cv_work.notify_all();  // the next queued job may now be first in line
        return PT_OK;
    }
    if (j.state != PT_JOB_RUNNING) return set_error(PT_ERR_STATE, "job %llu has finished", (unsigned long long)id);
    // the worker stops scheduling it and releases it after its rounds in flight
    j.cancel = true;
    s->cv_work.notify_all();
    return PT_OK;
}

pt_status pt_server_stats_get(pt_server* s, pt_server_stats* out) {
    if (!s || !out) return set_error(PT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->bytes_in_use = s->bytes_in_use;
    if (!s->in_flight.empty()) out->busy_s += now_s() - s->busy_since;
    return PT_OK;
}

}  // extern "C"
