#pragma once
#include "internal.h"

namespace vr {
int render_device(vr_scene* s, const vr_render_desc* desc, float* d_rgba, vr_render_stats* stats, hipStream_t st);
// Completes the counters of the last render after the stream has been synchronised.
void render_finish_stats(vr_scene* s, vr_render_stats* stats);
int tonemap_device(const vr_post_desc* desc, const float* d_src, float* d_dst, hipStream_t st);
}  // namespace vr
