#pragma once
#include "internal.h"

namespace vr {
// One request: a batch of one.
int render_device(vr_scene* s, const vr_render_desc* desc, float* d_rgba, hipStream_t st);
// Radiance-cache role of one request of a batch.
struct CacheReq {
    uint32_t cmode;            // 0 none, 1 cached
    uint32_t cache_depth;      // query vertex
    uint32_t train_vertices;   // recorded vertices per training path
    uint32_t train_threshold;  // fraction of training paths * 2^32
    uint32_t frame;
    float query_roughness;     // query at the first vertex from cache_depth on with roughness input >= this
};
// Plans, uploads and enqueues a batch (optionally captured into a graph and launched).
int render_batch(vr_ctx* ctx, const vr_request* items, uint32_t n, const vr_tuning* tuning, uint32_t flags,
                 hipStream_t st, const CacheReq* creq = nullptr, vr_cache* cache = nullptr);
// One frame of the cached renderer.
int render_cached(vr_scene* s, vr_cache* c, const vr_render_desc* desc, const vr_cache_frame* f, float* d_rgba,
                  hipStream_t st);
// Launches the last captured batch again.
int relaunch_batch(vr_ctx* ctx, hipStream_t st);
// Counters of the last render or batch on the context (waits for its read-back).
int render_finish_stats(vr_ctx* ctx, vr_render_stats* stats);
// Completion time of each request of the last time-stamped batch, ms after its start.
int batch_latencies(vr_ctx* ctx, double* ms, uint32_t n);
int autotune_batch(vr_ctx* ctx, const vr_request* items, uint32_t n, uint32_t flags, uint32_t reps,
                   hipStream_t st, vr_tuning* best, float* best_ms);
void free_render_state(vr_ctx* ctx);
int tonemap_device(const vr_post_desc* desc, const float* d_src, float* d_dst, hipStream_t st);
}  // namespace vr
