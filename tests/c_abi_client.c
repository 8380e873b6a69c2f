/* A C host calling the library through include/voidray_hip.h only, the way a
 * Rust -sys crate or any other native caller binds it: build a scene (two
 * triangles of a quad at z = 1, a Lambertian material, a uniform sky), trace
 * rays with known answers, render a few samples, tonemap.  Prints one line per
 * check and "ok" at the end; exits nonzero at the first failure. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "voidray_hip.h"

#define CHECK(call)                                                              \
    do {                                                                         \
        int rc_ = (call);                                                        \
        if (rc_ != VR_OK) {                                                      \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, vr_last_error()); \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(void) {
    vr_ctx* ctx = NULL;
    CHECK(vr_create(0, &ctx));

    /* quad [0,1]^2 at z = 1 as two triangles (Surfaces::quad's index order) */
    const float pos[12] = {0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1};
    const uint32_t idx[6] = {0, 1, 2, 2, 0, 3};
    vr_mesh mesh = {4, 2, pos, NULL, NULL, idx};
    const uint32_t surface_mesh[1] = {0};
    vr_object obj = {0, 0};
    vr_material mat = {VR_MAT_LAMBERTIAN, {0.5f, 0.5f, 0.5f}, 0.0f, -1, -1, 1.5f, 1.0f, 0.0f, 0.0f, 0};
    vr_scene_desc sd = {0};
    sd.num_meshes = 1;
    sd.meshes = &mesh;
    sd.num_surfaces = 1;
    sd.surface_mesh = surface_mesh;
    sd.num_objects = 1;
    sd.objects = &obj;
    sd.num_materials = 1;
    sd.materials = &mat;
    sd.environment.kind = VR_ENV_UNIFORM;
    sd.environment.color[0] = sd.environment.color[1] = sd.environment.color[2] = 1.0f;
    vr_camera cam = {{0.5f, 0.5f, -1.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 1.0f, 0.0f}, 0.8f, 0, 0.0f, {0, 0, 0}};
    sd.camera = cam;
    vr_scene* scene = NULL;
    CHECK(vr_scene_upload(ctx, &sd, &scene));

    /* rays along +z: (0.75, 0.25) crosses triangle 0 (below the diagonal),
     * (0.25, 0.75) triangle 1, (2, 2) misses; all at t = 1 from z = 0 */
    const float o[9] = {0.75f, 0.25f, 0.0f, 0.25f, 0.75f, 0.0f, 2.0f, 2.0f, 0.0f};
    const float d[9] = {0, 0, 1, 0, 0, 1, 0, 0, 1};
    uint32_t prim[3];
    float t[3], u[3], v[3];
    vr_trace_desc td = {3, 1e-5f, VR_TRACE_CLOSEST, VR_ENGINE_WIDE};
    CHECK(vr_trace(scene, &td, o, d, prim, t, u, v, NULL));
    printf("trace: prim %u %u %u, t %.6f %.6f\n", prim[0], prim[1], prim[2], t[0], t[1]);
    if (prim[0] != 0 || prim[1] != 1 || prim[2] != 0xFFFFFFFFu || t[0] != 1.0f || t[1] != 1.0f) return 2;
    uint8_t hit[3];
    td.mode = VR_TRACE_ANY;
    CHECK(vr_trace(scene, &td, o, d, NULL, NULL, NULL, NULL, hit));
    printf("any hit: %u %u %u\n", hit[0], hit[1], hit[2]);
    if (!hit[0] || !hit[1] || hit[2]) return 3;

    /* a 16x16 render at 4 spp, then the ACES tonemap */
    enum { W = 16, H = 16 };
    float* rgba = calloc(W * H * 4, sizeof(float));
    float* out = calloc(W * H * 4, sizeof(float));
    vr_render_desc rd = {W, H, 4, 4, 0, 8, 3.0f, VR_MODE_FULL, VR_RENDER_ALPHA_PER_CALL, 0, 7, 0, VR_ENGINE_WIDE};
    vr_render_stats st = {0};
    CHECK(vr_render(scene, &rd, rgba, &st));
    double sum = 0.0;
    for (int i = 0; i < W * H * 4; ++i) {
        if (!isfinite(rgba[i])) return 4;
        sum += rgba[i];
    }
    printf("render: %llu samples, %llu segments, sum %.4f, alpha %.1f\n", (unsigned long long)st.samples,
           (unsigned long long)st.segments, sum, rgba[3]);
    if (st.samples != (uint64_t)W * H * 4 || sum <= 0.0 || rgba[3] != 1.0f) return 5;
    vr_post_desc pd = {W, H, 1.0f, 2.2f, 0.0f, VR_TONEMAP_ACES, 0, 0};
    CHECK(vr_tonemap(ctx, &pd, rgba, out));
    for (int i = 0; i < W * H * 4; ++i)
        if (!isfinite(out[i])) return 6;
    printf("tonemap: first pixel %.4f %.4f %.4f %.1f\n", out[0], out[1], out[2], out[3]);

    free(rgba);
    free(out);
    CHECK(vr_scene_release(scene));
    CHECK(vr_destroy(ctx));
    printf("ok\n");
    return 0;
}
