// C ABI entry points (include/voidray_hip.h).  Every function returns an int
// status and never lets an exception cross the boundary; the message of the
// last failure on this thread is in vr_last_error().
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "internal.h"
#include "traverse.h"
#include "render.h"

namespace vr {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {
struct HostStaging {
    void* p = nullptr;
    size_t bytes = 0;
};
thread_local HostStaging g_dev_stage;

int ensure_dev_stage(size_t bytes) {
    if (g_dev_stage.bytes >= bytes) return 0;
    if (g_dev_stage.p) (void)hipFree(g_dev_stage.p);
    g_dev_stage.p = nullptr;
    g_dev_stage.bytes = 0;
    if (hipMalloc(&g_dev_stage.p, bytes) != hipSuccess) return fail(VR_ERR_OOM, "staging allocation failed");
    g_dev_stage.bytes = bytes;
    return 0;
}

template <class F> int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(VR_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(VR_ERR_INVALID, e.what());
    } catch (...) {
        return fail(VR_ERR_INVALID, "unknown exception");
    }
}
}  // namespace
}  // namespace vr

using namespace vr;

extern "C" {

const char* vr_last_error(void) { return g_last_error.c_str(); }

int vr_create(int device, vr_ctx** out) {
    return guarded([&] {
        if (!out) return fail(VR_ERR_INVALID, "null output pointer");
        int count = 0;
        VR_HIP_CHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) return fail(VR_ERR_INVALID, "device index out of range");
        VR_HIP_CHECK(hipSetDevice(device));
        auto* c = new vr_ctx();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return fail(VR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        // set-up of the no-allocation trace scratch on this device (a failure
        // here is reported again by vr_trace_device_noalloc_async)
        void* scratch_ptr = nullptr;
        (void)noalloc_scratch(device, &scratch_ptr);
        set_error("");
        *out = c;
        return 0;
    });
}

int vr_destroy(vr_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();  // batches may run on caller streams
    free_render_state(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

void* vr_stream(vr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int vr_scene_upload(vr_ctx* ctx, const vr_scene_desc* desc, vr_scene** out) {
    return guarded([&] {
        if (!ctx || !out) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(ctx->device));
        auto* s = new vr_scene();
        int rc = build_scene(ctx, desc, s);
        if (rc) {
            free_scene(s);
            delete s;
            return rc;
        }
        *out = s;
        return 0;
    });
}

int vr_scene_release(vr_scene* s) {
    if (!s) return 0;
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    free_scene(s);
    if (s->host_counters) (void)hipHostFree(s->host_counters);
    delete s;
    return 0;
}

int vr_scene_get_stats(const vr_scene* s, vr_scene_stats* out) {
    if (!s || !out) return fail(VR_ERR_INVALID, "null argument");
    *out = s->stats;
    return 0;
}

int vr_trace_device_async(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d,
                          uint32_t* d_prim, float* d_t, float* d_u, float* d_v, uint8_t* d_hit, void* stream) {
    return guarded([&] {
        if (!s || !desc) return fail(VR_ERR_INVALID, "null argument");
        hipStream_t st = stream ? (hipStream_t)stream : s->ctx->stream;
        return trace_device(s, desc, d_o, d_d, d_prim, d_t, d_u, d_v, d_hit, st);
    });
}

int vr_trace_device_noalloc_async(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d,
                                  uint32_t* d_prim, float* d_t, float* d_u, float* d_v, uint8_t* d_hit,
                                  void* stream) {
    return guarded([&] {
        if (!s || !desc) return fail(VR_ERR_INVALID, "null argument");
        hipStream_t st = stream ? (hipStream_t)stream : s->ctx->stream;
        return trace_device_noalloc(s, desc, d_o, d_d, d_prim, d_t, d_u, d_v, d_hit, st);
    });
}

int vr_trace(vr_scene* s, const vr_trace_desc* desc, const float* o, const float* d, uint32_t* prim, float* t,
             float* u, float* v, uint8_t* hit) {
    return guarded([&] {
        if (!s || !desc) return fail(VR_ERR_INVALID, "null argument");
        uint32_t n = desc->num_rays;
        if (n == 0) return 0;
        if (!o || !d) return fail(VR_ERR_INVALID, "null ray buffers");
        hipStream_t st = s->ctx->stream;
        size_t rb = 3 * sizeof(float) * (size_t)n;
        size_t off_d = rb, off_prim = 2 * rb, off_t = off_prim + 4 * (size_t)n, off_u = off_t + 4 * (size_t)n,
               off_v = off_u + 4 * (size_t)n, off_hit = off_v + 4 * (size_t)n;
        size_t total = off_hit + (size_t)n + 256;
        int rc = ensure_dev_stage(total);
        if (rc) return rc;
        char* base = (char*)g_dev_stage.p;
        VR_HIP_CHECK(hipMemcpyAsync(base, o, rb, hipMemcpyHostToDevice, st));
        VR_HIP_CHECK(hipMemcpyAsync(base + off_d, d, rb, hipMemcpyHostToDevice, st));
        rc = trace_device(s, desc, (const float*)base, (const float*)(base + off_d),
                          prim ? (uint32_t*)(base + off_prim) : nullptr, t ? (float*)(base + off_t) : nullptr,
                          u ? (float*)(base + off_u) : nullptr, v ? (float*)(base + off_v) : nullptr,
                          hit ? (uint8_t*)(base + off_hit) : nullptr, st);
        if (rc) return rc;
        if (prim) VR_HIP_CHECK(hipMemcpyAsync(prim, base + off_prim, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        if (t) VR_HIP_CHECK(hipMemcpyAsync(t, base + off_t, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        if (u) VR_HIP_CHECK(hipMemcpyAsync(u, base + off_u, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        if (v) VR_HIP_CHECK(hipMemcpyAsync(v, base + off_v, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        if (hit) VR_HIP_CHECK(hipMemcpyAsync(hit, base + off_hit, (size_t)n, hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

int vr_trace_host(const vr_scene_desc* sd, const vr_trace_desc* desc, const float* o, const float* d,
                  uint32_t* prim, float* t, float* u, float* v, uint8_t* hit, vr_scene_stats* stats,
                  uint32_t* fallbacks) {
    return guarded([&] {
        if (!sd || !desc) return fail(VR_ERR_INVALID, "null argument");
        const int32_t mode = desc->mode & ~VR_TRACE_TIMING;  // timing: device calls only
        if (mode != VR_TRACE_CLOSEST && mode != VR_TRACE_ANY) return fail(VR_ERR_INVALID, "bad trace mode");
        if (desc->engine != VR_ENGINE_WIDE && desc->engine != VR_ENGINE_REFERENCE &&
            desc->engine != VR_ENGINE_WATERTIGHT)
            return fail(VR_ERR_INVALID, "bad engine");
        const uint32_t n = desc->num_rays;
        if (n && (!o || !d)) return fail(VR_ERR_INVALID, "null ray buffers");
        vr::HostScene H;
        int rc = vr::build_host_scene(sd, H);
        if (rc) return rc;
        if (stats) {
            *stats = H.stats;
            stats->num_tri_records = H.stats.num_tri_records;
        }
        const vr::SceneView& sv = H.view;
        const bool any = mode == VR_TRACE_ANY;
        constexpr int kStack = 16, kSpill = 240;
        u32x2 lds[kStack], spill[kSpill];
        const vr::WideStack ws{lds, 1, spill, 1, kSpill};
        uint32_t fb = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const vf3 oi = mk3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), di = mk3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
            vr::HitOut h;
            const bool wt = desc->engine == VR_ENGINE_WATERTIGHT;
            bool exact = desc->engine == VR_ENGINE_REFERENCE;
            if (!exact) {
                // the kernels' template variant: [any + 2 * watertight + 4 * instances]
                int (*const wide[8])(const vr::SceneView&, vf3, vf3, float, float, const vr::WideStack&, vr::HitOut&) = {
                    vr::trace_wide<kStack, false, false, false>, vr::trace_wide<kStack, true, false, false>,
                    vr::trace_wide<kStack, false, true, false>,  vr::trace_wide<kStack, true, true, false>,
                    vr::trace_wide<kStack, false, false, true>,  vr::trace_wide<kStack, true, false, true>,
                    vr::trace_wide<kStack, false, true, true>,   vr::trace_wide<kStack, true, true, true>};
                const int r = wide[(any ? 1 : 0) + (wt ? 2 : 0) + (sv.num_instances ? 4 : 0)](sv, oi, di, desc->t_min,
                                                                                             kInf, ws, h);
                if (r == vr::kStepFallback) {
                    exact = true;
                    ++fb;
                }
            }
            // any hit: t_max = inf, so any hit counts
            if (exact) wt ? vr::trace_exact<true>(sv, oi, di, desc->t_min, h) : vr::trace_exact(sv, oi, di, desc->t_min, h);
            if (prim) prim[i] = h.prim;
            if (t) t[i] = h.t;
            if (u) u[i] = h.u;
            if (v) v[i] = h.v;
            if (hit) hit[i] = h.prim != 0xFFFFFFFFu ? 1 : 0;
        }
        if (fallbacks) *fallbacks = fb;
        return 0;
    });
}

int vr_trace_last_fallbacks(vr_scene* s, uint32_t* count) {
    if (!s || !count) return fail(VR_ERR_INVALID, "null argument");
    VR_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    *count = s->host_counters ? s->host_counters[0] : 0;
    return 0;
}

int vr_trace_last_timing(vr_scene* s, float* wide_ms) {
    if (!s || !wide_ms) return fail(VR_ERR_INVALID, "null argument");
    if (!s->trace_timed || !s->trace_ev[1])
        return fail(VR_ERR_INVALID, "the last trace was not a timed wide-engine trace (mode | VR_TRACE_TIMING)");
    VR_HIP_CHECK(hipEventSynchronize(s->trace_ev[1]));
    VR_HIP_CHECK(hipEventElapsedTime(wide_ms, s->trace_ev[0], s->trace_ev[1]));
    return 0;
}

int vr_render_device_async(vr_scene* s, const vr_render_desc* desc, float* d_rgba, vr_render_stats* stats,
                           void* stream) {
    return guarded([&] {
        if (!s || !desc || !d_rgba) return fail(VR_ERR_INVALID, "null argument");
        (void)stats;
        hipStream_t st = stream ? (hipStream_t)stream : s->ctx->stream;
        return render_device(s, desc, d_rgba, st);
    });
}

int vr_render(vr_scene* s, const vr_render_desc* desc, float* rgba, vr_render_stats* stats) {
    return guarded([&] {
        if (!s || !desc || !rgba) return fail(VR_ERR_INVALID, "null argument");
        size_t bytes = 4 * sizeof(float) * (size_t)desc->width * desc->height;
        hipStream_t st = s->ctx->stream;
        int rc = ensure_dev_stage(bytes);
        if (rc) return rc;
        VR_HIP_CHECK(hipMemcpyAsync(g_dev_stage.p, rgba, bytes, hipMemcpyHostToDevice, st));
        hipEvent_t e0, e1;
        VR_HIP_CHECK(hipEventCreate(&e0));
        VR_HIP_CHECK(hipEventCreate(&e1));
        VR_HIP_CHECK(hipEventRecord(e0, st));
        rc = render_device(s, desc, (float*)g_dev_stage.p, st);
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        VR_HIP_CHECK(hipEventRecord(e1, st));
        VR_HIP_CHECK(hipMemcpyAsync(rgba, g_dev_stage.p, bytes, hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        float ms = 0.f;
        VR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (stats) {
            stats->kernel_ms = ms;
            return render_finish_stats(s->ctx, stats);
        }
        return 0;
    });
}

int vr_render_last_stats(vr_scene* s, vr_render_stats* stats) {
    if (!s || !stats) return fail(VR_ERR_INVALID, "null argument");
    stats->kernel_ms = 0.0;
    return render_finish_stats(s->ctx, stats);
}

int vr_render_batch_async(vr_ctx* ctx, const vr_request* requests, uint32_t count, const vr_tuning* tuning,
                          uint32_t flags, void* stream) {
    return guarded([&] {
        if (!ctx || (count && !requests)) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        return render_batch(ctx, requests, count, tuning, flags, st);
    });
}

int vr_render_cached_async(vr_scene* s, vr_cache* cache, const vr_render_desc* desc, const vr_cache_frame* frame,
                           float* d_rgba, void* stream) {
    return guarded([&] {
        if (!s || !cache || !desc || !frame || !d_rgba) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : s->ctx->stream;
        return render_cached(s, cache, desc, frame, d_rgba, st);
    });
}

int vr_batch_relaunch_async(vr_ctx* ctx, void* stream) {
    return guarded([&] {
        if (!ctx) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(ctx->device));
        return relaunch_batch(ctx, stream ? (hipStream_t)stream : ctx->stream);
    });
}

int vr_batch_latencies(vr_ctx* ctx, double* ms, uint32_t count) {
    return guarded([&] {
        if (!ctx || (count && !ms)) return fail(VR_ERR_INVALID, "null argument");
        return batch_latencies(ctx, ms, count);
    });
}

int vr_batch_last_stats(vr_ctx* ctx, vr_render_stats* stats) {
    return guarded([&] {
        if (!ctx || !stats) return fail(VR_ERR_INVALID, "null argument");
        stats->kernel_ms = 0.0;
        return render_finish_stats(ctx, stats);
    });
}

int vr_batch_autotune(vr_ctx* ctx, const vr_request* requests, uint32_t count, uint32_t flags, uint32_t reps,
                      void* stream, vr_tuning* best, float* best_ms) {
    return guarded([&] {
        if (!ctx || !best || (count && !requests)) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(ctx->device));
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        return autotune_batch(ctx, requests, count, flags, reps, st, best, best_ms);
    });
}

int vr_tonemap_device_async(vr_ctx* ctx, const vr_post_desc* desc, const float* d_src, float* d_dst, void* stream) {
    return guarded([&] {
        if (!ctx || !desc || !d_src || !d_dst) return fail(VR_ERR_INVALID, "null argument");
        hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
        return tonemap_device(desc, d_src, d_dst, st);
    });
}

int vr_tonemap(vr_ctx* ctx, const vr_post_desc* desc, const float* src, float* dst) {
    return guarded([&] {
        if (!ctx || !desc || !src || !dst) return fail(VR_ERR_INVALID, "null argument");
        size_t bytes = 4 * sizeof(float) * (size_t)desc->width * desc->height;
        int rc = ensure_dev_stage(2 * bytes + 256);
        if (rc) return rc;
        char* base = (char*)g_dev_stage.p;
        hipStream_t st = ctx->stream;
        VR_HIP_CHECK(hipMemcpyAsync(base, src, bytes, hipMemcpyHostToDevice, st));
        // dst starts as a copy of itself so that pixels outside a clipped dispatch keep their value
        VR_HIP_CHECK(hipMemcpyAsync(base + bytes, dst, bytes, hipMemcpyHostToDevice, st));
        rc = tonemap_device(desc, (const float*)base, (float*)(base + bytes), st);
        if (rc) return rc;
        VR_HIP_CHECK(hipMemcpyAsync(dst, base + bytes, bytes, hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipStreamSynchronize(st));
        return 0;
    });
}

}  // extern "C"
