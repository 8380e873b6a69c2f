// Tonemapping pass: the HIP counterpart of the Vulkan compute shader
// voidray_renderer/src/shaders/post_process.glsl:23-49 with the curves of
// shaders/tonemapping.glsl:2-40.  One thread per pixel over the whole image;
// clip_to_1024 reproduces the fixed [1024/8, 1024/8] dispatch of
// render/post_process.rs:74 (pixels outside it are left untouched).  Modes from
// VR_TONEMAP_LEGACY_NONE on are the legacy viewer's fragment pass instead.
// pow and exp are the f32-only routines of f32math.h (the device library's
// powf goes through f64 internally).
#include "f32math.h"
#include "render.h"

namespace vr {

__device__ __forceinline__ void tm_aces(float& r, float& g, float& b) {
    // GLSL mat3(...) is column-major: the first three values are column 0.
    float x = 0.59719f * r + 0.35458f * g + 0.04823f * b;
    float y = 0.076f * r + 0.90834f * g + 0.01566f * b;
    float z = 0.0284f * r + 0.13383f * g + 0.83777f * b;
    x = (x * (x + 0.0245786f) - 0.000090537f) / (x * (0.983729f * x + 0.432951f) + 0.238081f);
    y = (y * (y + 0.0245786f) - 0.000090537f) / (y * (0.983729f * y + 0.432951f) + 0.238081f);
    z = (z * (z + 0.0245786f) - 0.000090537f) / (z * (0.983729f * z + 0.432951f) + 0.238081f);
    r = 1.60475f * x + -0.53108f * y + -0.07367f * z;
    g = -0.10208f * x + 1.10813f * y + -0.00605f * z;
    b = -0.00327f * x + -0.07276f * y + 1.07602f * z;
}

__device__ __forceinline__ void tm_reinhard(float& r, float& g, float& b) {
    const float white = 2.f;
    float luma = r * 0.2126f + g * 0.7152f + b * 0.0722f;
    float tl = luma * (1.f + luma / (white * white)) / (1.f + luma);
    float k = tl / luma;
    r *= k;
    g *= k;
    b *= k;
}

__device__ __forceinline__ float tm_filmic1(float c) {
    float x = fmaxf(0.f, c - 0.004f);
    return (x * (6.2f * x + .5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
}

__device__ __forceinline__ float tm_u2(float c) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f, W = 11.2f;
    float white = ((W * (A * W + C * B) + D * E) / (W * (A * W + B) + D * F)) - E / F;
    float x = c * 2.f;
    x = ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
    return x / white;
}

// Curves only the legacy viewer has (voidray/src/shaders/tonemapping.glsl:17-45).
__device__ __forceinline__ float tm_simple_reinhard1(float c) {
    const float exposure = 1.5f;
    return c * (exposure / (1.f + c / exposure));
}

__device__ __forceinline__ void tm_luma_reinhard(float& r, float& g, float& b) {
    float luma = r * 0.2126f + g * 0.7152f + b * 0.0722f;
    float k = (luma / (1.f + luma)) / luma;
    r *= k;
    g *= k;
    b *= k;
}

__device__ __forceinline__ float tm_rombindahouse1(float c) { return f32m::exp_cr(-1.0f / (2.72f * c + 0.15f)); }

// voidray/src/shaders/viewport_fragment.glsl:27-67: sampled = texel * scale,
// colour = sampled.rgb * exposure, curve, pow(1 / gamma) (Filmic: gamma 1).
__device__ __forceinline__ f32x4 legacy_post(const vr_post_desc& p, f32x4 s) {
    float r = (s.x * p.scale) * p.exposure, g = (s.y * p.scale) * p.exposure, b = (s.z * p.scale) * p.exposure;
    float gamma = p.gamma;
    switch (p.tonemap) {
        case VR_TONEMAP_LEGACY_SIMPLE_ACES: tm_aces(r, g, b); break;
        case VR_TONEMAP_LEGACY_SIMPLE_REINHARD:
            r = tm_simple_reinhard1(r);
            g = tm_simple_reinhard1(g);
            b = tm_simple_reinhard1(b);
            break;
        case VR_TONEMAP_LEGACY_LUMA_REINHARD: tm_luma_reinhard(r, g, b); break;
        case VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD: tm_reinhard(r, g, b); break;
        case VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE:
            r = tm_rombindahouse1(r);
            g = tm_rombindahouse1(g);
            b = tm_rombindahouse1(b);
            break;
        case VR_TONEMAP_LEGACY_FILMIC:
            r = tm_filmic1(r);
            g = tm_filmic1(g);
            b = tm_filmic1(b);
            gamma = 1.0f;
            break;
        case VR_TONEMAP_LEGACY_UNCHARTED2:
            r = tm_u2(r);
            g = tm_u2(g);
            b = tm_u2(b);
            break;
        default: break;
    }
    const float ig = 1.0f / gamma;
    return f32x4{f32m::pow_cr(r, ig), f32m::pow_cr(g, ig), f32m::pow_cr(b, ig), p.keep_alpha ? s.w * p.scale : 1.f};
}

__global__ __launch_bounds__(256) void k_tonemap(vr_post_desc p, const f32x4* __restrict__ src, f32x4* __restrict__ dst) {
    uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
    uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= p.width || y >= p.height) return;
    if (p.clip_to_1024 && (x >= 1024 || y >= 1024)) return;
    size_t i = (size_t)y * p.width + x;
    f32x4 s = src[i];
    if (p.tonemap >= VR_TONEMAP_LEGACY_NONE) {
        dst[i] = legacy_post(p, s);
        return;
    }
    float ex = f32m::pow_cr(2.f, p.exposure);
    float r = (s.x * p.scale) * ex, g = (s.y * p.scale) * ex, b = (s.z * p.scale) * ex;
    switch (p.tonemap) {
        case VR_TONEMAP_ACES: tm_aces(r, g, b); break;
        case VR_TONEMAP_REINHARD: tm_reinhard(r, g, b); break;
        case VR_TONEMAP_FILMIC:
            r = tm_filmic1(r);
            g = tm_filmic1(g);
            b = tm_filmic1(b);
            break;
        case VR_TONEMAP_UNCHARTED2:
            r = tm_u2(r);
            g = tm_u2(g);
            b = tm_u2(b);
            break;
        default: break;
    }
    float ig = 1.f / p.gamma;
    dst[i] = f32x4{f32m::pow_cr(r, ig), f32m::pow_cr(g, ig), f32m::pow_cr(b, ig), 1.f};
}

int tonemap_device(const vr_post_desc* desc, const float* d_src, float* d_dst, hipStream_t st) {
    if (desc->width == 0 || desc->height == 0) return 0;
    if (((uintptr_t)d_src & 15) || ((uintptr_t)d_dst & 15)) return fail(VR_ERR_INVALID, "buffers must be 16-byte aligned");
    if (desc->tonemap < VR_TONEMAP_NONE || (desc->tonemap > VR_TONEMAP_UNCHARTED2 && desc->tonemap < VR_TONEMAP_LEGACY_NONE) ||
        desc->tonemap > VR_TONEMAP_LEGACY_UNCHARTED2)
        return fail(VR_ERR_INVALID, "unknown tonemap");
    dim3 grid((desc->width + 15) / 16, (desc->height + 15) / 16);
    hipLaunchKernelGGL(k_tonemap, grid, dim3(256), 0, st, *desc, (const f32x4*)d_src, (f32x4*)d_dst);
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace vr
