// tracer.cpp - pt_tracer: the PathTracer of the reference (path_tracer.rs:59-358).
//
// A tracer owns the tile's accumulation buffer (float4 per tile pixel, the
// running sum of samples) and storage image (EXPOSURE * sum / samples), and
// the per-sample parameters: image size, tile origin, max bounces and the
// 1-based sample count, which set_tile / set_max_bounces / scene updates reset
// exactly as PathTracer does (path_tracer.rs:264-313, 272-282).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "pt_runtime.h"
#include "tracer.h"

using namespace ptcore;

namespace ptcore {

void tracer_sync_version(pt_tracer* t) {
    if (t->scene->version != t->scene_version) {
        t->scene_version = t->scene->version;
        t->params.sample_count = 0;  // reset_accumulation (path_tracer.rs:278-280)
    }
}

pt_status tracer_require_tile(const pt_tracer* t) {
    if (!t->accum || !t->storage)
        return set_error(PT_ERR_STATE, "tracer has no tile buffers (a failed set_tile released them; set a tile again)");
    return PT_OK;
}

HostJob tracer_job(pt_tracer* t, uint32_t samples) {
    HostJob hj;
    hj.scene = t->scene;
    ptd::WaveJob& j = hj.job;
    j.scene = 0;
    j.image_w = t->params.image_size[0];
    j.image_h = t->params.image_size[1];
    j.origin_x = t->params.tile_origin[0];
    j.origin_y = t->params.tile_origin[1];
    j.tile_w = t->tile_w;
    j.tile_h = t->tile_h;
    j.max_bounces = t->params.max_bounces;
    j.first_sample = t->params.sample_count + 1u;
    j.samples = samples;
    j.accum = t->accum;
    j.storage = t->storage;
    return hj;
}

}  // namespace ptcore

namespace {

// (Re)allocate the tile buffers. On failure the tracer is left with none and
// a zero tile, so no later launch can address a buffer of another size.
pt_status alloc_tile(pt_tracer* t, uint32_t w, uint32_t h) {
    (void)hipFree(t->accum);
    (void)hipFree(t->storage);
    t->accum = t->storage = nullptr;
    t->tile_w = t->tile_h = 0;
    size_t n = (size_t)w * h * sizeof(float4);
    hipError_t e = hipMalloc(&t->accum, n);
    if (e == hipSuccess) e = hipMalloc(&t->storage, n);
    if (e == hipSuccess) e = hipMemsetAsync(t->accum, 0, n, t->ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->storage, 0, n, t->ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(t->accum);
        (void)hipFree(t->storage);
        t->accum = t->storage = nullptr;
        return hip_error(e, "tile buffers");
    }
    t->tile_w = w;
    t->tile_h = h;
    return PT_OK;
}

// device staging of at least `bytes`, kept across calls
pt_status tracer_scratch(pt_tracer* t, size_t bytes, void** out) {
    if (bytes > t->scratch_bytes) {
        hipError_t se = hipStreamSynchronize(t->ctx->stream);
        if (se != hipSuccess) return hip_error(se, "tracer staging");
        (void)hipFree(t->scratch);
        t->scratch = nullptr;
        t->scratch_bytes = 0;
        hipError_t e = hipMalloc(&t->scratch, bytes);
        if (e != hipSuccess) {
            t->scratch = nullptr;
            return hip_error(e, "tracer staging");
        }
        t->scratch_bytes = bytes;
    }
    *out = t->scratch;
    return PT_OK;
}

// PathTracer::validate_tile_size (path_tracer.rs:361-382): nonzero, and the
// accumulation buffer must be addressable
pt_status validate_tile(uint32_t w, uint32_t h) {
    if (w == 0 || h == 0) return set_error(PT_ERR_INVALID, "tile size must be nonzero");
    if ((uint64_t)w * h > (1ull << 31)) return set_error(PT_ERR_INVALID, "tile of %ux%u is too large", w, h);
    return PT_OK;
}

void sync_version(pt_tracer* t) { tracer_sync_version(t); }

ptd::TileArgs tile_args(const pt_tracer* t, uint32_t first_sample, uint32_t samples) {
    ptd::TileArgs a;
    a.image_w = t->params.image_size[0];
    a.image_h = t->params.image_size[1];
    a.origin_x = t->params.tile_origin[0];
    a.origin_y = t->params.tile_origin[1];
    // only the part of the tile within the image is traced (path_tracer.rs:336-340)
    a.tile_w = t->tile_w;
    a.tile_h = t->tile_h;
    a.max_bounces = t->params.max_bounces;
    a.first_sample = first_sample;
    a.samples = samples;
    return a;
}

pt_status trace_samples(pt_tracer* t, uint32_t samples) {
    Context* c = t->ctx;
    if (pt_status ts = tracer_require_tile(t)) return ts;
    if (pt_status dst = use_device(c->device)) return dst;
    sync_version(t);
    const uint32_t first = t->params.sample_count + 1u;
    ptd::TileArgs a = tile_args(t, first, samples);
    pt_status st;
    if (t->schedule == PT_SCHEDULE_MEGAKERNEL) {
        const uint32_t blocks = ((t->tile_w + 15u) / 16u) * ((t->tile_h + 15u) / 16u);
        const uint32_t grid = std::min<uint32_t>(blocks, c->cu_count * 8u);
        if ((st = c->ensure_overflow(t->scene->bvh_depth + 1u, grid * 256u))) return st;
        const bool counting = t->profiling >= 2;
        if (counting) {
            hipError_t me = hipMemsetAsync(c->counters, 0, 16 * sizeof(unsigned long long), c->stream);
            if (me != hipSuccess) return hip_error(me, "traversal counters");
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (t->profiling) {  // the megakernel's one launch, at every level
            t->log.take(e0, e1);
            if (e0 && hipEventRecord(e0, c->stream) != hipSuccess) e0 = nullptr;  // timing is best effort
        }
        hipError_t e = launch_megakernel(t->scene->view, a, t->accum, t->storage, c->ovf, grid,
                                         counting ? c->counters : nullptr, c->stream);
        if (e != hipSuccess) return hip_error(e, "megakernel launch");
        if (t->profiling) {
            if (e1 && hipEventRecord(e1, c->stream) != hipSuccess) e1 = nullptr;
            t->log.pairs.push_back({EV_TOTAL, e0, e1});
        }
        if (counting) {
            unsigned long long h[4];
            if ((e = hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
                (e = hipStreamSynchronize(c->stream)) != hipSuccess)
                return hip_error(e, "traversal counters");
            for (int k = 0; k < 4; ++k) t->counters[k] += h[k];
        }
    } else {
        if (!t->wave) {
            t->wave = new Wavefront();
            t->wave->ctx = c;
        }
        if ((st = t->wave->render({tracer_job(t, samples)}, t->profiling, &t->log, t->counters))) return st;
        if (t->profiling >= 2) {
            unsigned long long h[6];
            hipError_t e = hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_error(e, "traversal counters");
            for (int k = 2; k < 6; ++k) t->counters[k] += h[k];
        }
    }
    t->params.sample_count += samples;
    return PT_OK;
}

}  // namespace

extern "C" {

pt_status pt_tracer_create(pt_context* c, pt_scene* s, uint32_t tile_w, uint32_t tile_h, pt_tracer** out) {
    if (!c || !s || !out) return set_error(PT_ERR_INVALID, "pt_tracer_create: null argument");
    *out = nullptr;
    if (s->ctx != c) return set_error(PT_ERR_INVALID, "scene belongs to another context");
    pt_status st = validate_tile(tile_w, tile_h);
    if (st) return st;
    if (pt_status dst = use_device(c->device)) return dst;
    pt_tracer* t = new pt_tracer();
    t->ctx = c;
    t->scene = s;
    t->scene_version = s->version;
    t->params.image_size[0] = tile_w;
    t->params.image_size[1] = tile_h;
    t->params.max_bounces = PT_DEFAULT_MAX_BOUNCES;
    if ((st = alloc_tile(t, tile_w, tile_h))) {
        pt_tracer_destroy(t);
        return st;
    }
    *out = t;
    return PT_OK;
}

pt_status pt_tracer_destroy(pt_tracer* t) {
    if (!t) return PT_OK;
    (void)hipSetDevice(t->ctx->device);
    const hipError_t pending = hipStreamSynchronize(t->ctx->stream);
    delete t->wave;
    (void)hipFree(t->accum);
    (void)hipFree(t->storage);
    (void)hipFree(t->scratch);
    delete t;
    return pending == hipSuccess ? PT_OK : hip_error(pending, "work pending at tracer destruction");
}

pt_status pt_tracer_set_schedule(pt_tracer* t, int schedule) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    if (schedule != PT_SCHEDULE_MEGAKERNEL && schedule != PT_SCHEDULE_WAVEFRONT)
        return set_error(PT_ERR_INVALID, "unknown schedule %d", schedule);
    t->schedule = schedule;
    return PT_OK;
}

pt_status pt_tracer_set_tile(pt_tracer* t, uint32_t iw, uint32_t ih, uint32_t ox, uint32_t oy, uint32_t tw, uint32_t th) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    if (iw == 0 || ih == 0) return set_error(PT_ERR_INVALID, "image size must be nonzero");
    if (ox >= iw || oy >= ih) return set_error(PT_ERR_INVALID, "tile origin must be within the image");
    pt_status st = validate_tile(tw, th);
    if (st) return st;
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    if (tw != t->tile_w || th != t->tile_h) {
        hipError_t se = hipStreamSynchronize(t->ctx->stream);
        if (se != hipSuccess) return hip_error(se, "tile resize");
        if ((st = alloc_tile(t, tw, th))) return st;
    }
    t->params.image_size[0] = iw;
    t->params.image_size[1] = ih;
    t->params.tile_origin[0] = ox;
    t->params.tile_origin[1] = oy;
    t->params.sample_count = 0;
    t->scene_version = t->scene->version;
    return PT_OK;
}

pt_status pt_tracer_set_max_bounces(pt_tracer* t, uint32_t mb) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    if (mb > PT_MAX_BOUNCES_LIMIT) return set_error(PT_ERR_INVALID, "max bounces %u above %u", mb, PT_MAX_BOUNCES_LIMIT);
    if (t->params.max_bounces != mb) {
        t->params.max_bounces = mb;
        t->params.sample_count = 0;
    }
    return PT_OK;
}

pt_status pt_tracer_set_sample_count(pt_tracer* t, uint32_t n) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    sync_version(t);
    t->params.sample_count = n;
    return PT_OK;
}

uint32_t pt_tracer_sample_count(const pt_tracer* t) { return t ? t->params.sample_count : 0; }
uint32_t pt_tracer_max_bounces(const pt_tracer* t) { return t ? t->params.max_bounces : 0; }

pt_status pt_tracer_trace(pt_tracer* t, uint32_t samples, uint32_t* traced) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    uint32_t n = std::min(samples, PT_MAX_SAMPLES_PER_TRACE);  // path_tracer.rs:322
    if (traced) *traced = 0;
    if (n == 0) return PT_OK;
    pt_status st = trace_samples(t, n);
    if (st == PT_OK && traced) *traced = n;
    return st;
}

pt_status pt_tracer_trace_all(pt_tracer* t, uint32_t samples) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    if (samples == 0) return PT_OK;
    return trace_samples(t, samples);
}

pt_status pt_tracer_read_accumulation(pt_tracer* t, float* rgba) {
    if (!t || !rgba) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status ts = tracer_require_tile(t)) return ts;
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    hipError_t e = hipMemcpyAsync(rgba, t->accum, (size_t)t->tile_w * t->tile_h * sizeof(float4), hipMemcpyDeviceToHost,
                                  t->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->ctx->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "read accumulation");
}

pt_status pt_tracer_write_accumulation(pt_tracer* t, const float* rgba) {
    if (!t || !rgba) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status ts = tracer_require_tile(t)) return ts;
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    hipError_t e = hipMemcpyAsync(t->accum, rgba, (size_t)t->tile_w * t->tile_h * sizeof(float4), hipMemcpyHostToDevice,
                                  t->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->ctx->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "write accumulation");
}

pt_status pt_tracer_read_storage(pt_tracer* t, float* rgba) {
    if (!t || !rgba) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status ts = tracer_require_tile(t)) return ts;
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    hipError_t e = hipMemcpyAsync(rgba, t->storage, (size_t)t->tile_w * t->tile_h * sizeof(float4), hipMemcpyDeviceToHost,
                                  t->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->ctx->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "read storage");
}

pt_status pt_tracer_display_rgba8(pt_tracer* t, uint8_t* rgba) {
    if (!t || !rgba) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status ts = tracer_require_tile(t)) return ts;
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    const uint32_t n = t->tile_w * t->tile_h;
    void* buf = nullptr;
    pt_status st = tracer_scratch(t, (size_t)n * 4, &buf);
    if (st) return st;
    hipError_t e = launch_display(t->storage, n, (uchar4*)buf, t->ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba, buf, (size_t)n * 4, hipMemcpyDeviceToHost, t->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->ctx->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "display");
}

pt_status pt_render_tile_noalloc(pt_scene* s, const PtTraceParameters* p, uint32_t tile_w, uint32_t tile_h,
                                 uint32_t samples, float* accum, float* storage) {
    if (!s || !p || !accum) return set_error(PT_ERR_INVALID, "pt_render_tile_noalloc: null argument");
    pt_status st = validate_tile(tile_w, tile_h);
    if (st) return st;
    const uint32_t iw = p->image_size[0], ih = p->image_size[1];
    if (iw == 0 || ih == 0) return set_error(PT_ERR_INVALID, "image size must be nonzero");
    if (p->tile_origin[0] >= iw || p->tile_origin[1] >= ih)
        return set_error(PT_ERR_INVALID, "tile origin must be within the image");
    if (p->max_bounces > PT_MAX_BOUNCES_LIMIT)
        return set_error(PT_ERR_INVALID, "max bounces %u above %u", p->max_bounces, PT_MAX_BOUNCES_LIMIT);
    if (p->sample_count == 0) return set_error(PT_ERR_INVALID, "sample_count is the 1-based index of the first sample");
    if ((uint64_t)p->sample_count + samples > 0xffffffffull) return set_error(PT_ERR_INVALID, "sample index overflow");
    if (samples == 0) return PT_OK;
    Context* c = s->ctx;
    // the megakernel's overflow stacks: depth beyond the LDS part x grid threads,
    // held to the static scratch array by the grid size
    const size_t levels = s->bvh_depth + 1u > ptd::MEGA_LDS_DEPTH ? s->bvh_depth + 1u - ptd::MEGA_LDS_DEPTH : 1u;
    const size_t per_block = levels * 256u * sizeof(uint32_t);
    const size_t max_grid = ptd::NOALLOC_SCRATCH_BYTES / per_block;
    if (max_grid == 0)
        return set_error(PT_ERR_INVALID, "BVH depth %u needs %zu bytes of traversal stack per block, over the %zu-byte scratch",
                         s->bvh_depth, per_block, ptd::NOALLOC_SCRATCH_BYTES);
    if (!c->noalloc_scratch) return set_error(PT_ERR_STATE, "no scratch address on this context's device");
    if (pt_status dst = use_device(c->device)) return dst;
    const uint32_t blocks = ((tile_w + 15u) / 16u) * ((tile_h + 15u) / 16u);
    const uint32_t grid = (uint32_t)std::min<size_t>(std::min<uint32_t>(blocks, c->cu_count * 8u), max_grid);
    ptd::TileArgs a;
    a.image_w = iw;
    a.image_h = ih;
    a.origin_x = p->tile_origin[0];
    a.origin_y = p->tile_origin[1];
    a.tile_w = tile_w;
    a.tile_h = tile_h;
    a.max_bounces = p->max_bounces;
    a.first_sample = p->sample_count;
    a.samples = samples;
    const hipError_t e = launch_megakernel(s->view, a, (float4*)accum, (float4*)storage, c->noalloc_scratch, grid,
                                           nullptr, c->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "no-alloc tile render");
}

pt_status pt_tonemap_rgba8(pt_context* c, const float* rgba, uint32_t n, uint8_t* out) {
    if (!c || !rgba || !out) return set_error(PT_ERR_INVALID, "null argument");
    if (n == 0) return PT_OK;
    if (pt_status dst = use_device(c->device)) return dst;
    DeviceBuffer buf;
    pt_status st = buf.alloc((size_t)n * (sizeof(float4) + 4));
    if (st) return st;
    float4* din = (float4*)buf.ptr;
    uchar4* dout = (uchar4*)(din + n);
    hipError_t e = hipMemcpyAsync(din, rgba, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_display(din, n, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "tone map");
}

pt_status pt_tracer_primary_hits(pt_tracer* t, uint32_t sample_count, int32_t* tri, float* tt) {
    if (!t || !tri || !tt) return set_error(PT_ERR_INVALID, "null argument");
    if (sample_count == 0) return set_error(PT_ERR_INVALID, "sample_count is 1-based");
    if (pt_status ts = tracer_require_tile(t)) return ts;
    Context* c = t->ctx;
    if (pt_status dst = use_device(c->device)) return dst;
    const uint32_t n = t->tile_w * t->tile_h;
    const uint32_t blocks = ((t->tile_w + 15u) / 16u) * ((t->tile_h + 15u) / 16u);
    const uint32_t grid = std::min<uint32_t>(blocks, c->cu_count * 8u);
    pt_status st = c->ensure_overflow(t->scene->bvh_depth + 1u, grid * 256u);
    if (st) return st;
    void* buf = nullptr;
    if ((st = tracer_scratch(t, (size_t)n * 8, &buf))) return st;
    int32_t* dtri = (int32_t*)buf;
    float* dt = (float*)(dtri + n);
    ptd::TileArgs a = tile_args(t, sample_count, 1);
    hipError_t e = launch_primary_hits(t->scene->view, a, dtri, dt, c->ovf, grid, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tri, dtri, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tt, dt, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? PT_OK : hip_error(e, "primary hits");
}

pt_status pt_tracer_set_profiling(pt_tracer* t, int level) {
    if (!t) return set_error(PT_ERR_INVALID, "null tracer");
    if (level < 0 || level > 3) return set_error(PT_ERR_INVALID, "profiling level %d not in 0..3", level);
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    hipError_t se = hipStreamSynchronize(t->ctx->stream);
    if (se != hipSuccess) return hip_error(se, "profiling switch");
    // level 3 times only the traversal launches (the megakernel's one kernel):
    // the fewest markers on the stream
    t->profiling = level == 3 ? 1 : level;
    t->log.kinds = level == 3 ? (1u << EV_EXTEND) : ~0u;
    t->log.reset();
    std::memset(t->counters, 0, sizeof(t->counters));
    return PT_OK;
}

pt_status pt_tracer_stats(pt_tracer* t, pt_trace_stats* out) {
    if (!t || !out) return set_error(PT_ERR_INVALID, "null argument");
    if (pt_status dst = use_device(t->ctx->device)) return dst;
    float ms[EV_KINDS];
    uint32_t launches[EV_KINDS];
    t->log.sum(ms, launches);
    std::memset(out, 0, sizeof(*out));
    out->closest_rays = t->counters[0];
    out->shadow_rays = t->counters[1];
    if (t->schedule == PT_SCHEDULE_MEGAKERNEL) {
        out->pair_tests = t->counters[2];
        out->tri_tests = t->counters[3];
    } else {
        out->pair_tests = t->counters[2];  // closest-hit traversal (extend)
        out->tri_tests = t->counters[3];
        out->shadow_pair_tests = t->counters[4];
        out->shadow_tri_tests = t->counters[5];
    }
    out->extend_launches = launches[EV_EXTEND];
    out->extend_ms = ms[EV_EXTEND];
    out->shade_ms = ms[EV_SHADE];
    out->shadow_ms = ms[EV_SHADOW];
    out->total_ms = ms[EV_TOTAL];
    out->extend_busy_ms = t->log.busy(EV_EXTEND);
    return PT_OK;
}

}  // extern "C"
