// Wavefront path tracer (filled in below).
#include "render.h"

namespace vr {
int render_device(vr_scene*, const vr_render_desc*, float*, vr_render_stats*, hipStream_t) {
    return fail(VR_ERR_UNSUPPORTED, "render not built yet");
}
void render_finish_stats(vr_scene*, vr_render_stats*) {}
}  // namespace vr
